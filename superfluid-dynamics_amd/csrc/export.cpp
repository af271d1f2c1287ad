// export.cpp -- the C ABI (include/cusuperhelium.h): drop-in CuSuperHelium entry points and the
// native sfh_* handle API, both on top of sfh::Solver.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/cusuperhelium.h"
#include "solver.hpp"

using sfh::Solver;

struct sfh_solver {
    std::unique_ptr<Solver> s;
};

namespace {

thread_local std::string g_last_error;

template <typename F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const sfh::ArgError& e) {
        g_last_error = e.what();
        return std::string(e.what()).rfind("unsupported node count", 0) == 0 ? -2 : -3;
    } catch (const sfh::NonFiniteError& e) {
        g_last_error = e.what();
        return -4;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return -1;
    } catch (...) {
        g_last_error = "unknown error";
        return -1;
    }
}

// adimensionalizeProperties (Export.cu:1222-1246): alpha = 3.5e-24 (constants.cuh:165),
// rho_He = 150.
struct Units {
    double rho, kappa, depth, base_length, base_time, base_energy;
};
Units adimensionalize(double L, double rho, double kappa, double depth) {
    const double alpha = 3.5e-24, rho_he = 150.0;
    Units u;
    u.base_length = L / (2.0 * sfh::kPi);
    const double base_acc = 3 * alpha / std::pow(depth, 4);
    u.base_time = std::sqrt(u.base_length / base_acc);
    u.base_energy = 3.0 * rho_he * alpha * std::pow(u.base_length, 4) / std::pow(depth, 4);
    const double stf = rho_he * u.base_length * u.base_length * u.base_length / (u.base_time * u.base_time);
    u.kappa = kappa / stf;
    u.depth = depth / u.base_length;
    u.rho = rho / rho_he;
    return u;
}

// HeliumBoundaryProblem as the reference exports build it (ProblemProperties defaults:
// infinite_depth = false, use_expansions = false, U = 0).
sfh::Phys helium_phys(const Units& u, bool infinite_depth, bool use_exp, int exp_order) {
    sfh::Phys ph{};
    ph.kind = 1;
    ph.image = infinite_depth ? 0 : 1;
    ph.use_exp = use_exp ? 1 : 0;
    ph.exp_order = exp_order;
    ph.rho_quirk = 1;
    ph.need_upper = 0;
    ph.rho = u.rho;
    ph.kappa = u.kappa;
    ph.depth = u.depth;
    ph.U = 0.0;
    ph.g = 1.0;
    return ph;
}

sfh::Phys phys_from(const sfh_props& p) {
    sfh::Phys ph{};
    ph.kind = p.kind;
    ph.image = (p.kind == SFH_HELIUM && !p.infinite_depth) ? 1 : 0;
    ph.use_exp = p.use_expansions;
    ph.exp_order = p.expansion_order;
    ph.rho_quirk = p.rho_quirk;
    ph.need_upper = (p.kind == SFH_WATER && p.rho != 0.0) ? 1 : 0;
    ph.rho = p.rho;
    ph.kappa = p.kappa;
    ph.depth = p.depth;
    ph.U = p.U;
    ph.g = p.g;
    return ph;
}

// Workspaces of the drop-in entry points, kept across calls (the reference rebuilds plans,
// handles and 11 buffers on every call, CuSuperHelium/Export.cu:194-265).
std::mutex g_cache_mu;
std::map<std::tuple<int, int, int, double, double, double, int, int, int>, std::unique_ptr<Solver>> g_cache;

Solver& cached_solver(const sfh::Phys& ph, int n, int batch) {
    auto key = std::make_tuple(n, batch, ph.kind, ph.rho, ph.kappa, ph.depth, ph.image, ph.use_exp, ph.exp_order);
    auto it = g_cache.find(key);
    if (it != g_cache.end()) return *it->second;
    if (g_cache.size() > 16) g_cache.clear();
    sfh::Options opt;
    auto s = std::make_unique<Solver>(ph, std::vector<int>(batch, n), opt);
    Solver& ref = *s;
    g_cache.emplace(key, std::move(s));
    return ref;
}

bool dropin_n(size_t N) {
    static const size_t ok[] = {1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 768, 1024, 1536, 2048, 4096};
    for (size_t v : ok)
        if (v == N) return true;
    return false;
}

int rhs_dropin(const double* x, const double* y, const double* phi, double* vx, double* vy, double* rhsPhi,
               double L, double rho, double kappa, double depth, size_t N, int batch) {
    if (!x || !y || !phi || !vx || !vy || !rhsPhi) {
        g_last_error = "null pointer argument";
        return -3;
    }
    return guarded([&] {
        const Units u = adimensionalize(L, rho, kappa, depth);
        std::lock_guard<std::mutex> lk(g_cache_mu);
        Solver& s = cached_solver(helium_phys(u, false, false, 1), (int)N, batch);
        s.set_state(x, y, phi);
        s.rhs(vx, vy, rhsPhi);
    });
}

}  // namespace

extern "C" {

int calculateRHSFromVectors(const double* x, const double* y, const double* phi, double* vx, double* vy,
                            double* rhsPhi, double L, double rho, double kappa, double depth, size_t N) {
    if (!dropin_n(N)) {
        g_last_error = "unsupported particle number " + std::to_string(N);
        return -2;
    }
    return rhs_dropin(x, y, phi, vx, vy, rhsPhi, L, rho, kappa, depth, N, 1);
}

int calculateRHS256FromVectors(const double* x, const double* y, const double* phi, double* vx, double* vy,
                               double* rhsPhi, double L, double rho, double kappa, double depth) {
    return rhs_dropin(x, y, phi, vx, vy, rhsPhi, L, rho, kappa, depth, 256, 1);
}

int calculateRHS2048FromVectors(const double* x, const double* y, const double* phi, double* vx, double* vy,
                                double* rhsPhi, double L, double rho, double kappa, double depth) {
    return rhs_dropin(x, y, phi, vx, vy, rhsPhi, L, rho, kappa, depth, 2048, 1);
}

int calculateRHS256FromVectorsBatched(const double* x, const double* y, const double* phi, double* vx,
                                      double* vy, double* rhsPhi, double L, double rho, double kappa,
                                      double depth, int batchSize) {
    static const int ok[] = {1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 768, 1024, 1536, 2048};
    bool good = false;
    for (int v : ok) good = good || v == batchSize;
    if (!good) {
        g_last_error = "unsupported batch size " + std::to_string(batchSize);
        return -2;
    }
    return rhs_dropin(x, y, phi, vx, vy, rhsPhi, L, rho, kappa, depth, 256, batchSize);
}

int calculateVorticities256FromVectors(const c_double* Z, const c_double* phi, double* a, c_double* Zp,
                                       c_double* Zpp, double L, double rho, double kappa, double depth) {
    if (!Z || !phi || !a) {
        g_last_error = "null pointer argument";
        return -3;
    }
    return guarded([&] {
        const int N = 256;
        std::vector<double> x(N), y(N), p(N), vx(N), vy(N), dp(N);
        for (int i = 0; i < N; ++i) {
            x[i] = Z[i].re;
            y[i] = Z[i].im;
            p[i] = phi[i].re;   // loadDataToDevice keeps the real part (SimulationRunner.cuh:244-260)
        }
        const Units u = adimensionalize(L, rho, kappa, depth);
        std::lock_guard<std::mutex> lk(g_cache_mu);
        Solver& s = cached_solver(helium_phys(u, false, false, 1), N, 1);
        s.set_state(x.data(), y.data(), p.data());
        std::vector<double2> zp(N), zpp(N);
        s.rhs(vx.data(), vy.data(), dp.data(), a, zp.data(), zpp.data());
        for (int i = 0; i < N; ++i) {
            if (Zp) Zp[i] = c_double{zp[i].x, zp[i].y};
            if (Zpp) Zpp[i] = c_double{zpp[i].x, zpp[i].y};
        }
    });
}

int calculateDerivativeFFT256(const c_double* input, c_double* output) {
    if (!input || !output) {
        g_last_error = "null pointer argument";
        return -3;
    }
    guarded([&] {
        const int N = 256;
        sfh::Phys ph{};
        ph.g = 1.0;
        std::lock_guard<std::mutex> lk(g_cache_mu);
        Solver& s = cached_solver(ph, N, 1);
        std::vector<double2> in(N), out(N);
        for (int i = 0; i < N; ++i) in[i] = make_double2(input[i].re, input[i].im);
        s.derivative(in.data(), out.data(), false, 1.0);
        for (int i = 0; i < N; ++i) output[i] = c_double{out[i].x, out[i].y};
    });
    return 0;   // the reference always returns 0 here (Export.cu:383)
}

int integrateSimulationRK4(double* initialState, double** statesOut, size_t* statesCount, double** timesOut,
                           size_t* timesCount, SimProperties* simProperties, RK4SolverOptions* rkOptions,
                           size_t N) {
    if (!initialState || !statesOut || !statesCount || !timesOut || !timesCount || !simProperties ||
        !rkOptions) {
        g_last_error = "null pointer argument";
        return -3;
    }
    if (!dropin_n(N)) {
        g_last_error = "unsupported particle number " + std::to_string(N);
        return -2;
    }
    return guarded([&] {
        const SimProperties sp = *simProperties;
        const Units u = adimensionalize(sp.L, sp.rho, sp.kappa, sp.depth);
        // adimensionalizeRK4SolverOptions (Export.cu:1213-1220)
        const double t0 = rkOptions->t0 / u.base_time;
        const double t1 = rkOptions->t1 / u.base_time;
        const double dt = rkOptions->timeStep / u.base_time;
        if (!std::isfinite(t0) || !std::isfinite(t1) || !std::isfinite(dt))
            throw sfh::ArgError("t0, t1 and the time step must be finite");
        if (!(dt > 0.0)) throw sfh::ArgError("time step must be positive");
        if (t1 < t0) throw sfh::ArgError("t1 must not precede t0");
        // runEvolution (AutonomousRungeKuttaStepper.cuh:421): floor((t1 - t0) / dt) steps
        const double nsteps = std::floor((t1 - t0) / dt);
        if (!(nsteps < 4.0e15)) throw sfh::ArgError("too many time steps");
        const size_t steps = (size_t)nsteps;
        const bool traj = rkOptions->returnTrajectory;
        const size_t nst = traj ? steps : 1;
        if (nst > SIZE_MAX / (sizeof(double) * 3 * N) - 1) throw sfh::ArgError("trajectory too large to allocate");
        sfh::Options opt;
        Solver s(helium_phys(u, sp.infinite_depth, sp.use_expansions, sp.expansion_order),
                 std::vector<int>(1, (int)N), opt);
        const double* x = initialState;
        s.set_state(x, x + N, x + 2 * N);
        double* states = (double*)std::malloc(sizeof(double) * 3 * N * (nst ? nst : 1));
        double* times = traj ? (double*)std::malloc(sizeof(double) * (steps ? steps : 1)) : nullptr;
        if (!states || (traj && !times)) {
            std::free(states);
            std::free(times);
            throw std::runtime_error("host allocation failed");
        }
        double t = t0;
        try {
            if (traj) {
                // every state through the device trajectory ring (copies overlap the stepping); the
                // times accumulate from t0 as the reference's currentTime += dt
                for (size_t done = 0; done < steps;) {
                    const int chunk = (int)std::min<size_t>(steps - done, 1u << 30);
                    s.step_record(dt, chunk, 1, states + 3 * N * done, times + done);
                    done += (size_t)chunk;
                }
                for (size_t k = 0; k < steps; ++k) {
                    t += dt;
                    times[k] = t;
                }
            } else {
                for (size_t done = 0; done < steps;) {   // Solver::step takes an int count
                    const int chunk = (int)std::min<size_t>(steps - done, 1u << 30);
                    s.step(dt, chunk);
                    done += (size_t)chunk;
                }
                s.get_state(states, states + N, states + 2 * N);
            }
        } catch (...) {
            std::free(states);
            std::free(times);
            throw;
        }
        *statesOut = states;
        *statesCount = nst;
        *timesOut = times;
        *timesCount = traj ? steps : 0;
    });
}

int integrateSimulationRK4_freeMemory(double* statesOut, double* timesOut) {
    std::free(statesOut);
    std::free(timesOut);
    return 0;
}

namespace {

// adimensionalizeOptomechanicalVariables (Export.cu:1248-1270) and
// get_current_intensity_drive_strength (LightIntensity.cuh:31-34), with the adimensional rho
sfh::Phys opto_drive(const Units& u, const COptomechanicalVariables& ov, bool augmented, bool timed) {
    const double hbar = 1.054571817e-34;   // constants.cuh:10
    sfh::Phys d{};
    d.drive = 1;
    d.augmented = augmented ? 1 : 0;
    d.timed = timed ? 1 : 0;
    d.gamma = ov.gamma * u.base_time;
    d.detuning = ov.detuning * u.base_time;
    d.G = ov.G * u.base_time * u.base_length;
    d.tau = ov.tau / u.base_time;
    d.x0 = ov.location_x0_mode / u.base_length;
    d.sigma = ov.sigma_optical_mode / u.base_length;
    d.max_intensity = ov.max_intensity;
    const double hbar_adim = hbar / u.base_energy / u.base_time;
    d.beta = ov.beta * (hbar_adim * d.G / d.tau / (std::pow(d.sigma, 2.0) * u.rho));
    d.damping = ov.damping_strength;
    d.strength = hbar / (u.base_energy * u.base_time * u.rho) * d.G / std::pow(d.sigma, 2.0);
    return d;
}

bool opto_n(size_t N) { return N >= 32 && N <= 8192 && (N & (N - 1)) == 0; }   // Export.cu:950-968,1186-1207

}  // namespace

int integrateAugmentedOptomechanicalSimulationRK4(double* initialState, double** statesOut, size_t* statesCount,
                                                  double** timesOut, size_t* timesCount, SimProperties* simProperties,
                                                  RK4SolverOptions* rkOptions, COptomechanicalVariables* ov, size_t N) {
    if (!initialState || !statesOut || !statesCount || !timesOut || !timesCount || !simProperties || !rkOptions ||
        !ov) {
        g_last_error = "null pointer argument";
        return -3;
    }
    if (!opto_n(N)) {   // the sizes Export.cu:1076-1093 dispatches
        g_last_error = "unsupported particle number " + std::to_string(N);
        return -2;
    }
    return guarded([&] {
        const SimProperties sp = *simProperties;
        const Units u = adimensionalize(sp.L, sp.rho, sp.kappa, sp.depth);
        const double t0 = rkOptions->t0 / u.base_time;
        const double t1 = rkOptions->t1 / u.base_time;
        const double dt = rkOptions->timeStep / u.base_time;
        if (!std::isfinite(t0) || !std::isfinite(t1) || !std::isfinite(dt))
            throw sfh::ArgError("t0, t1 and the time step must be finite");
        if (!(dt > 0.0)) throw sfh::ArgError("time step must be positive");
        if (t1 < t0) throw sfh::ArgError("t1 must not precede t0");
        const double nsteps = std::floor((t1 - t0) / dt);
        if (!(nsteps < 2.0e9)) throw sfh::ArgError("too many time steps");
        const size_t steps = (size_t)nsteps;
        if (steps > SIZE_MAX / (sizeof(double) * 4 * N) - 1) throw sfh::ArgError("trajectory too large to allocate");
        const sfh::Phys d = opto_drive(u, *ov, true, false);
        sfh::Options opt;
        Solver s(helium_phys(u, sp.infinite_depth, sp.use_expansions, sp.expansion_order),
                 std::vector<int>(1, (int)N), opt);
        s.set_drive(d);
        const double* x = initialState;
        s.set_state(x, x + N, x + 2 * N);
        s.set_delayed(x + 3 * N);
        double* states = (double*)std::malloc(sizeof(double) * 4 * N * (steps ? steps : 1));
        double* times = (double*)std::malloc(sizeof(double) * (steps ? steps : 1));
        if (!states || !times) {
            std::free(states);
            std::free(times);
            throw std::runtime_error("host allocation failed");
        }
        try {
            for (size_t done = 0; done < steps;) {
                const int chunk = (int)std::min<size_t>(steps - done, 1u << 30);
                s.step_record(dt, chunk, 1, states + 4 * N * done, times + done);
                done += (size_t)chunk;
            }
        } catch (...) {
            std::free(states);
            std::free(times);
            throw;
        }
        double t = t0;   // runEvolution's currentTime, logged after each step
        for (size_t k = 0; k < steps; ++k) {
            t += dt;
            times[k] = t;
            // Export.cu:1057 writes the phi entry into the fourth block
            std::memcpy(states + 4 * N * k + 3 * N, states + 4 * N * k + 2 * N, N * sizeof(double));
        }
        *statesOut = states;
        *statesCount = steps;
        *timesOut = times;
        *timesCount = steps;
    });
}

int calculateJacobian(const double* state, double* jac, double L, double rho, double kappa, double depth,
                      double epsilon, size_t N) {
    if (!state || !jac) {
        g_last_error = "null pointer argument";
        return -3;
    }
    if (N < 32 || N > 2048 || (N & (N - 1)) != 0) {   // the sizes Export.cu:495-518 dispatches
        g_last_error = "unsupported particle number " + std::to_string(N);
        return -2;
    }
    return guarded([&] {
        const Units u = adimensionalize(L, rho, kappa, depth);
        sfh::fd_jacobian(helium_phys(u, false, false, 1), (int)N, state, epsilon, jac);
    });
}

int integrateSimulationGL2(double* initialState, double** statesOut, size_t* statesCount, double** timesOut,
                           size_t* timesCount, SimProperties* simProperties, GaussLegendreOptions* go, size_t N) {
    if (!initialState || !statesOut || !statesCount || !timesOut || !timesCount || !simProperties || !go) {
        g_last_error = "null pointer argument";
        return -3;
    }
    if (N < 32 || N > 1024 || (N & (N - 1)) != 0) {   // the sizes Export.cu:705-725 dispatches
        g_last_error = "unsupported particle number " + std::to_string(N);
        return -2;
    }
    return guarded([&] {
        const SimProperties sp = *simProperties;
        const Units u = adimensionalize(sp.L, sp.rho, sp.kappa, sp.depth);
        sfh::Gl2Options o;
        o.tol = go->newtonTolerance;
        o.max_newton = (int)std::min<size_t>(go->maxNewtonIterations, 1000);
        o.max_halves = (int)std::min<size_t>(go->maxStepsHalves, 60);
        o.simplified_fallback = go->allowSimplifiedFallback;
        o.armijo_c = go->armijo_c;
        o.backtrack = go->backtrack;
        o.min_alpha = go->minAlpha;
        if (!(o.backtrack > 0.0 && o.backtrack < 1.0) || !(o.min_alpha > 0.0))
            throw sfh::ArgError("backtrack must lie in (0, 1) and minAlpha be positive");
        std::vector<double> s(initialState, initialState + 3 * N), traj, times;
        const bool want = go->returnTrajectory;
        sfh::gl2_integrate(helium_phys(u, sp.infinite_depth, sp.use_expansions, sp.expansion_order), (int)N, s.data(),
                           go->t0, go->t1, go->stepSize, o, want ? &traj : nullptr, want ? &times : nullptr);
        *statesOut = nullptr;
        *timesOut = nullptr;
        *statesCount = *timesCount = 0;
        if (want) {
            double* st = (double*)std::malloc(traj.size() * sizeof(double));
            double* tm = (double*)std::malloc(times.size() * sizeof(double));
            if (!st || !tm) {
                std::free(st);
                std::free(tm);
                throw std::runtime_error("host allocation failed");
            }
            std::memcpy(st, traj.data(), traj.size() * sizeof(double));
            std::memcpy(tm, times.data(), times.size() * sizeof(double));
            *statesOut = st;
            *timesOut = tm;
            *statesCount = *timesCount = times.size();
        }
    });
}

int integrateSimulationGL2_freeMemory(double* statesOut, double* timesOut) {
    std::free(statesOut);
    std::free(timesOut);
    return 0;
}

int sfh_jacobian(const sfh_props* props, int n, const double* state, double eps, double* jac) {
    if (!props || !state || !jac) {
        g_last_error = "null pointer argument";
        return -3;
    }
    return guarded([&] { sfh::fd_jacobian(phys_from(*props), n, state, eps, jac); });
}

int integrateAugmentedOptomechanicalSimulationRK4_freeMemory(double* statesOut, double* timesOut) {
    std::free(statesOut);
    std::free(timesOut);
    return 0;
}

int integrateOptomechanicalSimulationRK4(double* initialState, double** statesOut, size_t* statesCount,
                                         double** timesOut, size_t* timesCount, SimProperties* simProperties,
                                         RK4SolverOptions* rkOptions, COptomechanicalVariables* ov, size_t N) {
    if (!initialState || !statesOut || !statesCount || !timesOut || !timesCount || !simProperties || !rkOptions ||
        !ov) {
        g_last_error = "null pointer argument";
        return -3;
    }
    if (!opto_n(N)) {
        g_last_error = "unsupported particle number " + std::to_string(N);
        return -2;
    }
    return guarded([&] {
        const SimProperties sp = *simProperties;
        const Units u = adimensionalize(sp.L, sp.rho, sp.kappa, sp.depth);
        const double t0 = rkOptions->t0 / u.base_time;
        const double t1 = rkOptions->t1 / u.base_time;
        const double dt = rkOptions->timeStep / u.base_time;
        if (!std::isfinite(t0) || !std::isfinite(t1) || !std::isfinite(dt))
            throw sfh::ArgError("t0, t1 and the time step must be finite");
        if (!(dt > 0.0)) throw sfh::ArgError("time step must be positive");
        if (t1 < t0) throw sfh::ArgError("t1 must not precede t0");
        // runEvolution (RK4_Time_Dependent.cuh:386-412): (t1 - t0) / dt steps, truncated
        const double nsteps = std::floor((t1 - t0) / dt);
        if (!(nsteps < 2.0e9)) throw sfh::ArgError("too many time steps");
        const size_t steps = (size_t)nsteps;
        const bool traj = rkOptions->returnTrajectory;
        // the reference collects states and times only when the trajectory is returned
        const size_t nst = traj ? steps : 0;
        if (nst > SIZE_MAX / (sizeof(double) * 3 * N) - 1) throw sfh::ArgError("trajectory too large to allocate");
        sfh::Options opt;
        Solver s(helium_phys(u, sp.infinite_depth, sp.use_expansions, sp.expansion_order),
                 std::vector<int>(1, (int)N), opt);
        s.set_drive(opto_drive(u, *ov, false, true));
        const double* x = initialState;
        s.set_state(x, x + N, x + 2 * N);
        s.set_time(t0);   // setStartingTime(t0): the delayed intensity's clock starts at t0
        double* states = (double*)std::malloc(sizeof(double) * 3 * N * (nst ? nst : 1));
        double* times = (double*)std::malloc(sizeof(double) * (nst ? nst : 1));
        if (!states || !times) {
            std::free(states);
            std::free(times);
            throw std::runtime_error("host allocation failed");
        }
        try {
            if (traj) {
                for (size_t done = 0; done < steps;) {
                    const int chunk = (int)std::min<size_t>(steps - done, 1u << 30);
                    s.step_record(dt, chunk, 1, states + 3 * N * done, times + done);
                    done += (size_t)chunk;
                }
                // devTimes receives currentTime before it advances: the time at the start of the step
                double t = t0;
                for (size_t k = 0; k < steps; ++k) {
                    times[k] = t;
                    t += dt;
                }
            } else {
                for (size_t done = 0; done < steps;) {
                    const int chunk = (int)std::min<size_t>(steps - done, 1u << 30);
                    s.step(dt, chunk);
                    done += (size_t)chunk;
                }
                s.sync();
            }
        } catch (...) {
            std::free(states);
            std::free(times);
            throw;
        }
        *statesOut = states;
        *statesCount = nst;
        *timesOut = times;
        *timesCount = nst;
    });
}

int integrateOptomechanicalSimulationRK4_freeMemory(double* statesOut, double* timesOut) {
    std::free(statesOut);
    std::free(timesOut);
    return 0;
}

int calculateRhsAugmentedOptomechanical(double* state, double* rhs, SimProperties* simProperties,
                                        COptomechanicalVariables* ov, size_t N) {
    if (!state || !rhs || !simProperties || !ov) {
        g_last_error = "null pointer argument";
        return -3;
    }
    if (!opto_n(N)) {
        g_last_error = "unsupported particle number " + std::to_string(N);
        return -2;
    }
    return guarded([&] {
        const SimProperties sp = *simProperties;
        const Units u = adimensionalize(sp.L, sp.rho, sp.kappa, sp.depth);
        sfh::Options opt;
        Solver s(helium_phys(u, sp.infinite_depth, sp.use_expansions, sp.expansion_order),
                 std::vector<int>(1, (int)N), opt);
        s.set_drive(opto_drive(u, *ov, true, false));
        s.set_state(state, state + N, state + 2 * N);
        s.set_delayed(state + 3 * N);
        s.rhs(rhs, rhs + N, rhs + 2 * N);
        s.get_delayed_rhs(rhs + 3 * N);
    });
}

int calculatePerturbedStates256(const double* x, const double* y, const double* phi, c_double* Zperturbed, double L,
                                double rho, double kappa, double depth, double epsilon) {
    (void)L;
    (void)rho;
    (void)kappa;
    (void)depth;   // the reference takes them and does not use them (Export.cu:601-660)
    if (!x || !y || !phi || !Zperturbed) {
        g_last_error = "null pointer argument";
        return -3;
    }
    return guarded([&] { sfh::perturbed_states(x, y, phi, 256, epsilon, reinterpret_cast<double*>(Zperturbed)); });
}

void sfh_default_options(sfh_options* opt) {
    if (!opt) return;
    std::memset(opt, 0, sizeof(*opt));
    opt->precision = SFH_FP64;
    opt->solver = SFH_SOLVE_LU;
    opt->use_graph = 1;
    opt->max_iter = 32;
    opt->restart = 0;
    opt->force_trim = 0;
    opt->tol = 1e-12;
}

void sfh_default_props(sfh_props* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->kind = SFH_WATER;
    p->expansion_order = 1;
    p->rho_quirk = 1;
    p->depth = 1.0;
    p->g = 1.0;
}

int sfh_create(const sfh_props* props, const int* n, int n_surfaces, const sfh_options* opt, sfh_handle* out) {
    if (!props || !n || n_surfaces <= 0 || !out) {
        g_last_error = "invalid argument";
        return -3;
    }
    *out = nullptr;
    return guarded([&] {
        sfh::Options o;
        if (opt) {
            o.precision = opt->precision;
            o.solver = opt->solver;
            o.use_graph = opt->use_graph;
            o.max_iter = opt->max_iter;
            o.restart = opt->restart;
            o.force_trim = opt->force_trim;
            o.tol = opt->tol;
            o.pair_rows = opt->pair_rows;
        }
        const std::vector<int> sizes(n, n + n_surfaces);
        const sfh::Phys ph = phys_from(*props);
        if (o.solver == SFH_SOLVE_AUTO) {
            const bool mixed = o.precision == SFH_MIXED;
            int mx = 0;
            bool uniform = true;
            for (int v : sizes) {
                mx = std::max(mx, v);
                uniform = uniform && v == sizes[0];
            }
            if (uniform) {
                const sfh::Tuning t = sfh::tuned(ph, mx, mixed);
                o.solver = t.solver;
                o.precision = t.precision;
                o.pair_rows = t.pair_rows;
            } else if (mx <= sfh::kBatchMaxN) {
                o.solver = SFH_SOLVE_BATCHED;
            } else {
                const sfh::Tuning t = sfh::tuned(ph, mx, mixed);
                o.solver = SFH_SOLVE_GMRES;
                o.pair_rows = t.pair_rows;
            }
            if (o.solver == SFH_SOLVE_BATCHED) o.max_iter = std::min(o.max_iter, 16);
            if (o.solver != SFH_SOLVE_LU) o.max_iter = std::min(o.max_iter, sfh::kGmMaxIter);
        }
        auto h = std::make_unique<sfh_solver>();
        h->s = std::make_unique<Solver>(ph, sizes, o);
        *out = h.release();
    });
}

int sfh_destroy(sfh_handle h) {
    return guarded([&] { delete h; });
}

int sfh_total_nodes(sfh_handle h) { return h ? h->s->total_nodes() : -3; }

#define SFH_NEED(h)                              \
    if (!(h)) {                                  \
        g_last_error = "null handle";            \
        return -3;                               \
    }

int sfh_set_state(sfh_handle h, const double* x, const double* y, const double* phi) {
    SFH_NEED(h);
    return guarded([&] { h->s->set_state(x, y, phi); });
}

int sfh_get_state(sfh_handle h, double* x, double* y, double* phi) {
    SFH_NEED(h);
    return guarded([&] { h->s->get_state(x, y, phi); });
}

int sfh_rhs(sfh_handle h, double* vx, double* vy, double* dphi, double* mu, double* zp, double* zpp,
            double* qup) {
    SFH_NEED(h);
    return guarded([&] {
        h->s->rhs(vx, vy, dphi, mu, reinterpret_cast<double2*>(zp), reinterpret_cast<double2*>(zpp),
                  reinterpret_cast<double2*>(qup));
    });
}

int sfh_step(sfh_handle h, double dt, int nsteps) {
    SFH_NEED(h);
    return guarded([&] { h->s->step(dt, nsteps); });
}

int sfh_rhs_vjp(sfh_handle h, const double* gvx, const double* gvy, const double* gdphi, double* gx, double* gy,
                double* gphi) {
    SFH_NEED(h);
    if (!gvx || !gvy || !gdphi || !gx || !gy || !gphi) return -3;
    return guarded([&] { h->s->rhs_vjp(gvx, gvy, gdphi, gx, gy, gphi); });
}

int sfh_step_vjp(sfh_handle h, double dt, const double* gx_in, const double* gy_in, const double* gphi_in,
                 double* gx_out, double* gy_out, double* gphi_out) {
    SFH_NEED(h);
    if (!gx_in || !gy_in || !gphi_in || !gx_out || !gy_out || !gphi_out) return -3;
    return guarded([&] { h->s->step_vjp(dt, gx_in, gy_in, gphi_in, gx_out, gy_out, gphi_out); });
}

int sfh_rhs_vjp_augmented(sfh_handle h, const double* gvx, const double* gvy, const double* gdphi,
                          const double* gdD, double* gx, double* gy, double* gphi, double* gD) {
    SFH_NEED(h);
    if (!gvx || !gvy || !gdphi || !gdD || !gx || !gy || !gphi || !gD) return -3;
    return guarded([&] { h->s->rhs_vjp(gvx, gvy, gdphi, gx, gy, gphi, gdD, gD); });
}

int sfh_step_vjp_augmented(sfh_handle h, double dt, const double* gx_in, const double* gy_in, const double* gphi_in,
                           const double* gD_in, double* gx_out, double* gy_out, double* gphi_out, double* gD_out) {
    SFH_NEED(h);
    if (!gx_in || !gy_in || !gphi_in || !gD_in || !gx_out || !gy_out || !gphi_out || !gD_out) return -3;
    return guarded([&] { h->s->step_vjp(dt, gx_in, gy_in, gphi_in, gx_out, gy_out, gphi_out, gD_in, gD_out); });
}

int sfh_traj_forward(sfh_handle h, double dt, int steps, int checkpoint_every, int* record) {
    SFH_NEED(h);
    if (!record) return -3;
    return guarded([&] { *record = h->s->traj_forward(dt, steps, checkpoint_every); });
}

int sfh_traj_backward(sfh_handle h, int record, const double* gx_in, const double* gy_in, const double* gphi_in,
                      double* gx_out, double* gy_out, double* gphi_out) {
    SFH_NEED(h);
    if (!gx_in || !gy_in || !gphi_in || !gx_out || !gy_out || !gphi_out) return -3;
    return guarded([&] { h->s->traj_backward(record, gx_in, gy_in, gphi_in, gx_out, gy_out, gphi_out); });
}

int sfh_traj_release(sfh_handle h, int record) {
    SFH_NEED(h);
    return guarded([&] { h->s->traj_release(record); });
}

int sfh_get_mu(sfh_handle h, double* mu) {
    SFH_NEED(h);
    return guarded([&] { h->s->get_mu(mu); });
}

int sfh_sync(sfh_handle h) {
    SFH_NEED(h);
    return guarded([&] { h->s->sync(); });
}

int sfh_get_stats(sfh_handle h, sfh_stats* out) {
    SFH_NEED(h);
    if (!out) return -3;
    return guarded([&] {
        sfh::Stats st = h->s->stats();
        std::memset(out, 0, sizeof(*out));
        out->rhs_evals = st.rhs_evals;
        out->steps = st.steps;
        out->fallback_solves = st.fallback_solves;
        out->spin_timeouts = st.spin_timeouts;
        out->last_iterations = st.last_iterations;
        out->max_iterations = st.max_iterations;
        out->unconverged = st.unconverged;
        out->refinements = st.refinements;
        out->redone_steps = st.redone_steps;
        out->sequence_length = st.sequence_length;
        out->solver = h->s->solver();
        out->precision = h->s->precision();
        out->pair_rows = h->s->pair_rows();
    });
}

int sfh_rhs_noalloc(const sfh_props* props, int n, const double* x, const double* y, const double* phi, double* vx,
                    double* vy, double* dphi, double* mu, double tol, int max_iter, void* stream) {
    if (!props) {
        g_last_error = "null props";
        return -3;
    }
    return guarded([&] {
        sfh::rhs_noalloc(phys_from(*props), n, x, y, phi, vx, vy, dphi, mu, tol, max_iter,
                         static_cast<hipStream_t>(stream));
    });
}

size_t sfh_rhs_noalloc_scratch_bytes(void) { return sfh::noalloc_scratch_bytes(); }

static void tuning_out(const sfh::Tuning& t, sfh_tuning* out) {
    out->solver = t.solver;
    out->precision = t.precision;
    out->pair_rows = t.pair_rows;
    out->from_cache = t.from_cache;
    out->ms_per_step = t.ms_per_step;
}

int sfh_autotune(const sfh_props* props, int n, int allow_mixed, int steps, sfh_tuning* out) {
    if (!props || !out) return -3;
    return guarded([&] { tuning_out(sfh::autotune(phys_from(*props), n, allow_mixed != 0, steps), out); });
}

int sfh_tuned(const sfh_props* props, int n, int allow_mixed, sfh_tuning* out) {
    if (!props || !out) return -3;
    return guarded([&] { tuning_out(sfh::tuned(phys_from(*props), n, allow_mixed != 0), out); });
}

int sfh_set_drive(sfh_handle h, const sfh_drive* dr) {
    SFH_NEED(h);
    return guarded([&] {
        sfh::Phys d{};
        if (dr) {
            d.drive = 1;
            d.augmented = dr->augmented ? 1 : 0;
            d.timed = dr->timed ? 1 : 0;
            d.damping = dr->damping;
            d.strength = dr->strength;
            d.detuning = dr->detuning;
            d.gamma = dr->gamma;
            d.G = dr->G;
            d.max_intensity = dr->max_intensity;
            d.x0 = dr->x0;
            d.sigma = dr->sigma;
            d.beta = dr->beta;
            d.tau = dr->tau;
        }
        h->s->set_drive(d);
    });
}

int sfh_set_time(sfh_handle h, double t) {
    SFH_NEED(h);
    return guarded([&] { h->s->set_time(t); });
}

int sfh_set_delayed(sfh_handle h, const double* D) {
    SFH_NEED(h);
    if (!D) return -3;
    return guarded([&] { h->s->set_delayed(D); });
}

int sfh_get_delayed(sfh_handle h, double* D) {
    SFH_NEED(h);
    if (!D) return -3;
    return guarded([&] { h->s->get_delayed(D); });
}

int sfh_get_delayed_rhs(sfh_handle h, double* dD) {
    SFH_NEED(h);
    if (!dD) return -3;
    return guarded([&] { h->s->get_delayed_rhs(dD); });
}

int sfh_set_state_device(sfh_handle h, const void* state) {
    SFH_NEED(h);
    return guarded([&] { h->s->set_state_device(static_cast<const double2*>(state)); });
}

int sfh_get_state_device(sfh_handle h, void* state) {
    SFH_NEED(h);
    return guarded([&] { h->s->get_state_device(static_cast<double2*>(state)); });
}

int sfh_rhs_device(sfh_handle h, const void* state, void* rhs) {
    SFH_NEED(h);
    return guarded([&] { h->s->rhs_device(static_cast<const double2*>(state), static_cast<double2*>(rhs)); });
}

int sfh_step_record(sfh_handle h, double dt, int nsteps, int every, double* states, double* times) {
    SFH_NEED(h);
    return guarded([&] { h->s->step_record(dt, nsteps, every, states, times); });
}

int sfh_track_energies(sfh_handle h, int enable) {
    SFH_NEED(h);
    return guarded([&] { h->s->track_energies(enable != 0); });
}

int sfh_get_energies(sfh_handle h, double* out) {
    SFH_NEED(h);
    if (!out) return -3;
    return guarded([&] { h->s->energies(out); });
}

int sfh_device_state(sfh_handle h, void** Z, void** phi) {
    SFH_NEED(h);
    if (Z) *Z = h->s->device_Z();
    if (phi) *phi = h->s->device_phi();
    return 0;
}

const char* sfh_last_error(void) { return g_last_error.c_str(); }

int sfh_kernel_timing(sfh_handle h, int enable) {
    SFH_NEED(h);
    return guarded([&] { h->s->set_kernel_timing(enable != 0); });
}

int sfh_kernel_report(sfh_handle h, const char* tag, int* launches, double* ms, double* flops, double* bytes) {
    SFH_NEED(h);
    if (!tag) return -3;
    return guarded([&] {
        Solver::KernelReport r = h->s->kernel_report(tag);
        if (launches) *launches = r.launches;
        if (ms) *ms = r.ms;
        if (flops) *flops = r.flops;
        if (bytes) *bytes = r.bytes;
    });
}

int sfh_debug_lu_solve(int n, const double* A, const double* b, double* x, int* interchanged) {
    if (!A || !b || !x || n < 1) {
        g_last_error = "invalid argument";
        return -3;
    }
    return guarded([&] { sfh::debug_lu_solve(n, A, b, x, interchanged); });
}

}  // extern "C"
