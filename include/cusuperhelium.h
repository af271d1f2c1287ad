/*
 * cusuperhelium.h -- C ABI of the gfx950 (MI355X) Roberts boundary-integral RK4 stepper.
 *
 * Two groups of entry points:
 *
 * 1. Drop-in replacements for the reference CuSuperHelium DLL exports (paths relative to
 *    CuSuperHelium/CuSuperHelium/).  Same names, argument meaning, host-pointer semantics
 *    (caller-owned SoA float64 arrays in, caller-allocated arrays out, trajectories allocated
 *    with malloc and released by the *_freeMemory call) and SI-unit conversion
 *    (adimensionalizeProperties, Export.cu:1222-1246).  Like the reference they build the
 *    helium-film problem (HeliumBoundaryProblem).  Return codes: 0 success, -1 runtime/device
 *    error, -2 unsupported size (the reference falls off the end of the function here),
 *    -3 invalid argument (a NaN or infinite input state included), -4 a right-hand side came out
 *    NaN or infinite (coincident nodes make the dipole matrix singular, a trajectory blew up).
 *    No entry point ever calls exit().  The reference returns NaN states in these cases
 *    (ValueLogger.h:55-58 throws only when logging).
 *
 * 2. The native handle API (sfh_*): run-time sizes, many surfaces of different N per handle,
 *    nondimensional parameters, problem kind (water or helium film), state kept in HBM across
 *    calls, RK4 steps replayed from a captured hipGraph.
 *
 * Every call is synchronous with respect to host memory it reads or writes.
 */
#ifndef CUSUPERHELIUM_H
#define CUSUPERHELIUM_H

#include <stdbool.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------------------------
 * Reference C structs (ExportTypes.cuh:7-40), same field order and types.
 * ---------------------------------------------------------------------------------------- */
typedef struct __attribute__((aligned(16))) c_double {
    double re;
    double im;
} c_double;

typedef struct SimProperties {
    double L;
    double rho;
    double kappa;
    double depth;
    bool use_expansions;
    int expansion_order;
    bool infinite_depth;
} SimProperties;

typedef struct RK4SolverOptions {
    double timeStep;
    double t0;
    double t1;
    bool returnTrajectory;
} RK4SolverOptions;

/* ExportTypes.cuh:21-34 (times and step in the nondimensional units the reference passes through). */
typedef struct GaussLegendreOptions {
    double t0;
    double t1;
    double stepSize;
    double newtonTolerance;
    size_t maxNewtonIterations;
    bool allowSimplifiedFallback;
    bool returnTrajectory;
    double armijo_c;
    double backtrack;
    double minAlpha;
    size_t maxStepsHalves;
} GaussLegendreOptions;

/* ExportTypes.cuh:41-64, SI units (adimensionalizeOptomechanicalVariables, Export.cu:1248-1270). */
typedef struct COptomechanicalVariables {
    double detuning;
    double gamma;
    double G;
    double tau;
    double max_intensity;
    double initial_time;
    double location_x0_mode;
    double sigma_optical_mode;
    double beta;
    double damping_strength;
} COptomechanicalVariables;

/* ------------------------------------------------------------------------------------------
 * Drop-in entry points.
 * ---------------------------------------------------------------------------------------- */

/* Export.cuh:31 / Export.cu:560-599 -- RHS (vx + i vy = dZ/dt, rhsPhi = dPhi/dt) of one surface
 * of N nodes, N in {1,2,4,...,4096} plus 768 and 1536. */
int calculateRHSFromVectors(const double* x, const double* y, const double* phi, double* vx, double* vy,
                            double* rhsPhi, double L, double rho, double kappa, double depth, size_t N);

/* Export.cuh:30 / Export.cu:116-189 -- N = 256. */
int calculateRHS256FromVectors(const double* x, const double* y, const double* phi, double* vx, double* vy,
                               double* rhsPhi, double L, double rho, double kappa, double depth);

/* Export.cuh:33 / Export.cu:267-277 -- N = 2048. */
int calculateRHS2048FromVectors(const double* x, const double* y, const double* phi, double* vx, double* vy,
                                double* rhsPhi, double L, double rho, double kappa, double depth);

/* Export.cuh:51 / Export.cu:520-557 -- batchSize surfaces of N = 256 concatenated, batchSize in
 * {1,2,4,...,2048} plus 768 and 1536. */
int calculateRHS256FromVectorsBatched(const double* x, const double* y, const double* phi, double* vx,
                                      double* vy, double* rhsPhi, double L, double rho, double kappa,
                                      double depth, int batchSize);

/* Export.cuh:48 / Export.cu:279-349 -- dipole strengths a (the reference's "vorticities") and,
 * when the pointers are non-null, Z' and Z'' of a 256-node surface. */
int calculateVorticities256FromVectors(const c_double* Z, const c_double* phi, double* a, c_double* Zp,
                                       c_double* Zpp, double L, double rho, double kappa, double depth);

/* Export.cuh:49 / Export.cu:351-384 -- spectral first derivative (reference multiplier table,
 * unit scale) of a 256-sample complex field. */
int calculateDerivativeFFT256(const c_double* input, c_double* output);

/* Export.cuh:69 / Export.cu:748-770 (a stub returning 0 in the reference) -- classical RK4 from
 * t0 to t1 (SI seconds) with floor((t1 - t0)/dt) steps.  initialState = [x(N), y(N), phi(N)].
 * With returnTrajectory: *statesOut = malloc'd [steps][3N] states after every step and
 * *timesOut = malloc'd [steps] nondimensional times; otherwise one final state and no times.
 * N as for calculateRHSFromVectors. */
int integrateSimulationRK4(double* initialState, double** statesOut, size_t* statesCount, double** timesOut,
                           size_t* timesCount, SimProperties* simProperties, RK4SolverOptions* rkOptions,
                           size_t N);

/* Export.cuh:70 / Export.cu:772-777 -- releases the arrays returned by integrateSimulationRK4. */
int integrateSimulationRK4_freeMemory(double* statesOut, double* timesOut);

/* Export.cuh:50 / Export.cu:393-518 -- central-difference Jacobian (JacobianCalculator.cuh:223-284) of
 * the helium film right-hand side [vx, vy, dphi] with respect to the nondimensional state
 * state = [x(N), y(N), phi(N)] (SI L, rho, kappa, depth as for calculateRHSFromVectors):
 * jac[k * 3N + r] = d f_r / d state_k (column-major 3N x 3N), step epsilon.  N in {32, ..., 2048}.
 * The 6N perturbed surfaces are evaluated as batches of the FP64 LU path. */
int calculateJacobian(const double* state, double* jac, double L, double rho, double kappa, double depth,
                      double epsilon, size_t N);

/* Export.cuh:66 / Export.cu:663-740 -- two-stage Gauss-Legendre (GaussLegendre.cuh:190-560) of the
 * helium film from glCOptions->t0 to t1 with step stepSize (nondimensional, used as given, like the
 * reference), Newton on the stage slopes with the central-difference Jacobian (epsilon 1e-6), Armijo
 * backtracking, optional frozen-Jacobian fallback, step halving up to maxStepsHalves times.
 * initialState = [x(N), y(N), phi(N)].  With returnTrajectory: *statesOut = malloc'd [frames][3N]
 * starting with the initial state, *timesOut = malloc'd [frames]; without it both are NULL and the
 * counts 0, as in the reference.  N in {32, ..., 1024}.  -1 when a step fails at every size. */
int integrateSimulationGL2(double* initialState, double** statesOut, size_t* statesCount, double** timesOut,
                           size_t* timesCount, SimProperties* simProperties, GaussLegendreOptions* glCOptions,
                           size_t N);
/* Export.cuh:67 -- releases the arrays returned by integrateSimulationGL2. */
int integrateSimulationGL2_freeMemory(double* statesOut, double* timesOut);

/* Export.cuh:75 / Export.cu:981-1093 -- the optically driven helium film with the delayed intensity D
 * as a third state field (HeliumDrivenAutonomousProblem + AugmentedBoundaryIntegrator), classical RK4
 * from t0 to t1 (SI seconds), floor((t1 - t0)/dt) steps, every state logged.
 * initialState = [x(N), y(N), phi(N), D(N)]; *statesOut = malloc'd [steps][4N] laid out as the
 * reference writes it: x, y, phi and, in the fourth block, phi again (Export.cu:1057 copies the phi
 * entry there, not D); *timesOut = malloc'd [steps] nondimensional times.  N in {32, 64, ..., 8192}. */
int integrateAugmentedOptomechanicalSimulationRK4(double* initialState, double** statesOut, size_t* statesCount,
                                                  double** timesOut, size_t* timesCount, SimProperties* simProperties,
                                                  RK4SolverOptions* rkOptions,
                                                  COptomechanicalVariables* optomechanicalVariables, size_t N);
/* Export.cuh:76 -- releases the arrays returned by integrateAugmentedOptomechanicalSimulationRK4. */
int integrateAugmentedOptomechanicalSimulationRK4_freeMemory(double* statesOut, double* timesOut);

/* Export.cuh:72 / Export.cu:780-968 -- the optically driven helium film with the time-dependent
 * delayed intensity (HeliumWithOptomechanicalDrivingProblem + TimedBoundaryIntegrator + the RK4 of
 * RK4_Time_Dependent.cuh:170-412): stages at t_n, t_n + h/2, t_n + h/2, t_n + h; at a stage of time ts
 * node i takes D = I_i when ts equals the saved time, else a Dsaved_i + beta tau (1 - a) I_i with
 * a = exp(-(ts - saved) / tau); the first stage of each step saves D and its time (the clock starts at
 * t0); dPhi/dt += damping v + beta D + strength I.  initialState = [x(N), y(N), phi(N)].  With
 * returnTrajectory, *statesOut = malloc'd [steps][3N] (the state after each step) and *timesOut =
 * malloc'd [steps] nondimensional times of the start of each step, as the reference logs them
 * (RK4_Time_Dependent.cuh:402-408); without it both counts are 0, as in the reference.
 * Steps: (t1 - t0) / dt truncated.  N in {32, 64, ..., 8192}.  Differs from the reference where that
 * is racy: there every thread of the saving stage stores the saved time while others still read it
 * (DelayedIntensityTerm.cuh:20-37), so some nodes take D = I; here every node of a stage reads the
 * saved time of the previous step. */
int integrateOptomechanicalSimulationRK4(double* initialState, double** statesOut, size_t* statesCount,
                                         double** timesOut, size_t* timesCount, SimProperties* simProperties,
                                         RK4SolverOptions* rkOptions, COptomechanicalVariables* optomechanicalVariables,
                                         size_t N);
/* Export.cuh:73 -- releases the arrays returned by integrateOptomechanicalSimulationRK4. */
int integrateOptomechanicalSimulationRK4_freeMemory(double* statesOut, double* timesOut);

/* Export.cuh:79 / Export.cu:1108-1209 -- right-hand side of the augmented system (the drive of
 * integrateAugmentedOptomechanicalSimulationRK4) at state = [x(N), y(N), phi(N), D(N)]:
 * rhs = [vx(N), vy(N), dphi/dt(N), dD/dt(N)].  N in {32, 64, ..., 8192}.  Returns error codes where
 * the reference returns 0 after printing an exception. */
int calculateRhsAugmentedOptomechanical(double* state, double* rhs, SimProperties* simProperties,
                                        COptomechanicalVariables* optomechanicalVariables, size_t N);

/* Export.cuh:53 / Export.cu:601-660 -- the +epsilon perturbed states of the N = 256 surface
 * (createInitialBatchedZ, JacobianCalculator.cuh:11-76): Zperturbed holds 6 * 256^2 complex, batch
 * b = c 256 + p (c = 0: x, 1: y, 2: phi of node p perturbed) with its positions at [256 b, 256 b + 256)
 * and its potentials phi + 0i at [3 * 256^2 + 256 b, ...).  L, rho, kappa and depth are unused, as in
 * the reference. */
int calculatePerturbedStates256(const double* x, const double* y, const double* phi, c_double* Zperturbed, double L,
                                double rho, double kappa, double depth, double epsilon);

/* ------------------------------------------------------------------------------------------
 * Native handle API.
 * ---------------------------------------------------------------------------------------- */
typedef struct sfh_solver* sfh_handle;

enum { SFH_WATER = 0, SFH_HELIUM = 1 };
enum { SFH_FP64 = 0, SFH_MIXED = 1 };
/* LU: blocked FP64 LU (the reference's getrf/getrs semantics), one surface at a time or uniform
 * batches.  GMRES: FP64 GMRES over the stored matrix, all surfaces of the handle at once (every
 * N <= 4096); with SFH_MIXED the matrix is stored in FP32, r0 and a refinement residual are FP64.
 * BATCHED: one persistent launch per RK stage, one workgroup per surface (every N <= 1024), GMRES
 * inside the workgroup; with SFH_MIXED the matrix is FP32 and mu is refined in FP64 against a
 * matrix-free FP64 residual. */
enum { SFH_SOLVE_LU = 0, SFH_SOLVE_GMRES = 1, SFH_SOLVE_BATCHED = 2, SFH_SOLVE_AUTO = 3 };

/* Nondimensional problem (ProblemProperties.hpp:5-32 plus kind and g).  g scales the restoring
 * term of the Bernoulli equation (1 reproduces the reference).  rho_quirk = 1 reproduces the
 * V1[1] index of compute_rhs_phi_expression (createM.cuh:104), which matters only for rho != 0. */
typedef struct sfh_props {
    int kind;
    int infinite_depth;
    int use_expansions;
    int expansion_order;
    int rho_quirk;
    int reserved;
    double rho;
    double kappa;
    double depth;
    double U;
    double g;
} sfh_props;

typedef struct sfh_options {
    int precision;    /* SFH_FP64 (the drop-in arithmetic) or SFH_MIXED (SFH_SOLVE_BATCHED or _GMRES) */
    int solver;       /* SFH_SOLVE_LU, SFH_SOLVE_GMRES or SFH_SOLVE_BATCHED */
    int use_graph;    /* replay RK4 steps from a captured hipGraph */
    int max_iter;     /* SFH_SOLVE_GMRES: Arnoldi steps per solve, 1..32 (sfh_create rejects more: the
                         launch sequence has no restart); SFH_SOLVE_BATCHED: GMRES iteration cap,
                         rounded up to restart cycles of 8 */
    int restart;      /* unused (0): the GMRES of SFH_SOLVE_GMRES does not restart, the in-workgroup
                         GMRES of SFH_SOLVE_BATCHED restarts every 8 basis vectors */
    int force_trim;   /* test hook (0 = off): length of the trimmed GMRES launch sequence instead of
                         the calibrated one (at least 2), also for the warm-started transposed solves
                         of sfh_traj_backward; a chunk of steps (a segment of the sweep) in which a
                         solve runs out of launches is still run again with the length doubled */
    double tol;       /* iterative solver / refinement relative tolerance */
    int pair_rows;    /* rows per workgroup of the pair kernels (assembly, velocity): 1, 2 or 4; 0 = 4 */
    int reserved;
} sfh_options;

/* SFH_SOLVE_AUTO picks solver, precision (SFH_MIXED in the options allows the mixed forms) and
 * pair_rows from the autotuner: a single size (one surface or a uniform batch) takes the cached or
 * freshly measured sfh_tuned() result of that size; an uneven ensemble of surfaces of at most 1024
 * nodes takes SFH_SOLVE_BATCHED, a larger one SFH_SOLVE_GMRES with the tuned rows of its largest
 * size.  No counterpart in the reference, whose launch shapes and solver are fixed
 * (BoundaryProblem.cuh:27-34, MatrixSolver.cuh:113-172). */
typedef struct sfh_tuning {
    int solver;
    int precision;
    int pair_rows;
    int from_cache;       /* 1 when sfh_tuned found the result in the cache */
    double ms_per_step;   /* measured ms per RK4 step of the chosen form */
} sfh_tuning;
/* Times every candidate (LU and GMRES with 1, 2 and 4 rows per workgroup, the batched form for
 * n <= 1024; the mixed forms too when allow_mixed) for `steps` RK4 steps of a seeded surface of n nodes
 * after an 8-step warm-up, keeps the fastest whose solves converge, and caches it per (GPU
 * architecture, kind, image term, n, allow_mixed) in $SFH_TUNE_CACHE, $XDG_CACHE_HOME or
 * ~/.cache (cusuperhelium/tuning.txt). */
int sfh_autotune(const sfh_props* props, int n, int allow_mixed, int steps, sfh_tuning* out);
/* The cached result, or sfh_autotune(..., 10, ...) on a miss. */
int sfh_tuned(const sfh_props* props, int n, int allow_mixed, sfh_tuning* out);

typedef struct sfh_stats {
    long long rhs_evals;
    long long steps;
    int fallback_solves;   /* systems that needed row interchanges and went through GEPP */
    int spin_timeouts;
    int last_iterations;   /* GMRES: iterations of surface 0 in the last solve */
    int max_iterations;    /* GMRES: most iterations any surface needed in the last solve */
    int unconverged;       /* GMRES: solves of kept steps that reached the iteration cap (accumulated) */
    int refinements;       /* mixed precision, surface 0, last solve: FP64 refinement sweeps (batched);
                              1 if the refinement solve ran Arnoldi steps, else 0 (GMRES) */
    int redone_steps;      /* SFH_SOLVE_GMRES: steps run again because a solve of their chunk ran out of
                              launches of the trimmed sequence (accumulated) */
    int sequence_length;   /* SFH_SOLVE_GMRES: launches per solve of the trimmed step graph, 0 while the
                              warm-start history fills (then every solve gets max_iter + 1) */
    int solver;            /* the form the handle runs (what SFH_SOLVE_AUTO resolved to) */
    int precision;
    int pair_rows;
    int reserved;
} sfh_stats;

/* Defaults: FP64, LU, graph replay on, max_iter 32, tol 1e-12. */
void sfh_default_options(sfh_options* opt);
void sfh_default_props(sfh_props* props);

/* Creates a solver for n_surfaces independent periodic surfaces with n[s] nodes each
 * (n = 2^k or 3 * 2^k, 1 <= n <= 8192).  Node arrays of all calls are the surfaces'
 * node arrays concatenated in this order. */
int sfh_create(const sfh_props* props, const int* n, int n_surfaces, const sfh_options* opt, sfh_handle* out);
int sfh_destroy(sfh_handle h);
int sfh_total_nodes(sfh_handle h);
int sfh_set_state(sfh_handle h, const double* x, const double* y, const double* phi);
int sfh_get_state(sfh_handle h, double* x, double* y, double* phi);
/* RHS of the current state.  mu (dipole strengths), zp / zpp (interleaved re, im) and qup
 * (upper-fluid velocity, interleaved) are optional outputs (may be NULL). */
int sfh_rhs(sfh_handle h, double* vx, double* vy, double* dphi, double* mu, double* zp, double* zpp,
            double* qup);
/* Advances every surface by nsteps classical RK4 steps of size dt, in place.  The call may return
 * before the steps have run on the device; a stage slope that came out NaN or infinite is then
 * reported (-4) by the next call that waits for the device (sfh_get_state, sfh_rhs, sfh_get_stats,
 * sfh_sync, ...) until sfh_set_state gives the handle a new state.  sfh_set_state rejects a NaN or
 * infinite state (-3). */
int sfh_step(sfh_handle h, double dt, int nsteps);

/* Central-difference Jacobian of the right-hand side of one surface (nondimensional problem, n <= 2048,
 * n = 2^k or 3 * 2^k): jac[k * 3n + r] = d f_r / d s_k for s = [x, y, phi], f = [vx, vy, dphi]. */
int sfh_jacobian(const sfh_props* props, int n, const double* state, double eps, double* jac);

/* Optical drive, nondimensional (OptomechanicalVariables after adimensionalizeOptomechanicalVariables,
 * OptomechanicalVariables.h:4-27; strength = get_current_intensity_drive_strength, LightIntensity.cuh:31-34):
 *   dPhi/dt += damping v + strength I(y, x)            (HeliumDrivenAutonomousProblem, createM.cuh:138-149)
 *   I = gamma^2 max_intensity / 4 / ((detuning - G y)^2 + (gamma/2)^2) exp(-(x - x0)^2 / (2 sigma^2))
 * and with augmented = 1 the delayed intensity D as a third state field (AugmentedBoundaryIntegrator,
 * DelayedIntensityIntegrator.cuh:20-32):  dPhi/dt += D,  dD/dt = beta I - D / tau.
 * sfh_set_drive (NULL: drive off) zeroes D; sfh_set_state zeroes it too.  sfh_set_delayed /
 * sfh_get_delayed move D (total nodes); sfh_get_delayed_rhs returns dD/dt of the last sfh_rhs.
 * With the augmented drive, sfh_step_record frames are [x, y, phi, D] (4T).  Not with the adjoint
 * or sfh_rhs_device. */
typedef struct sfh_drive {
    int augmented;
    int timed;   /* 1: the time-dependent delayed intensity of integrateOptomechanicalSimulationRK4 (D
                    saved by the first stage of each step; sfh_set_time starts its clock); not with
                    augmented */
    double damping;
    double strength;
    double detuning;
    double gamma;
    double G;
    double max_intensity;
    double x0;
    double sigma;
    double beta;
    double tau;
} sfh_drive;
int sfh_set_drive(sfh_handle h, const sfh_drive* drive);
/* Sets the handle's time (0 after sfh_set_state); with the timed drive also the saved time of the
 * delayed intensity (TimedBoundaryIntegrator::setStartingTime, TimedBoundaryIntegrator.cuh:48-52). */
int sfh_set_time(sfh_handle h, double t);
int sfh_set_delayed(sfh_handle h, const double* D);
int sfh_get_delayed(sfh_handle h, double* D);
int sfh_get_delayed_rhs(sfh_handle h, double* dD);

/* Device-resident state in the reference's layout: complex128 [Z(T), Phi(T) + 0i] in device memory
 * (BaseBoundaryIntegrator.cuh:141-146; the imaginary part of Phi is ignored on input).
 * sfh_rhs_device is AutonomousProblem::run (AutonomousProblem.h:25, BaseBoundaryIntegrator.cuh:308-314):
 * rhs = [dZ/dt (T), dPhi/dt + 0i (T)] of `state`, leaving the handle's own state untouched.  Each
 * call has completed on return. */
int sfh_set_state_device(sfh_handle h, const void* state);
int sfh_get_state_device(sfh_handle h, void* state);
int sfh_rhs_device(sfh_handle h, const void* state, void* rhs);

/* Advances nsteps steps like sfh_step and records the state after every `every`-th step
 * (TrajectoryLogger::logTrajectory, TrajectoryLogger.cuh:65-73, whose device vector grows by one
 * state per logged step): frame f (f < nsteps / every) is states[f * 3T .. (f + 1) * 3T) =
 * [x(T), y(T), phi(T)] in the node order of sfh_get_state (T = sfh_total_nodes), and times[f] the
 * handle's time after that step (0 after sfh_set_state, advanced by dt per step).  Frames pass
 * through a device ring of up to 32 frames and are copied to host on a second stream while the
 * stepping continues; the call returns when every frame is in the caller's arrays. */
int sfh_step_record(sfh_handle h, double dt, int nsteps, int every, double* states, double* times);

/* Reverse mode of the right-hand side of water (any rho, the V1[1] quirk included) and of the
 * helium film (image terms, van der Waals term or its expansion, surface tension), with or without
 * the autonomous or augmented optical drive (the augmented one through the _augmented entry points
 * below; the time-dependent drive returns -3); handles of
 * SFH_SOLVE_BATCHED take every N <= 1024 (one workgroup per surface), LU and GMRES handles any N
 * (many workgroups per surface).  No counterpart in the reference, which has no adjoint.
 * sfh_rhs_vjp: cotangents (gvx, gvy, gdphi) of the right-hand side (vx, vy, dphi) at the current
 * state -> cotangents (gx, gy, gphi) of the state (x, y, phi).
 * sfh_step_vjp: cotangent of the state one RK4 step of size dt after the current state ->
 * cotangent of the current state; the handle's state is not changed.  Together with saved
 * checkpoints this gives the gradient of any function of a trajectory. */
int sfh_rhs_vjp(sfh_handle h, const double* gvx, const double* gvy, const double* gdphi, double* gx, double* gy,
                double* gphi);
int sfh_step_vjp(sfh_handle h, double dt, const double* gx_in, const double* gy_in, const double* gphi_in,
                 double* gx_out, double* gy_out, double* gphi_out);
/* The same for a handle with the augmented drive, whose state has the delayed intensity D as a third
 * field: gdD is the cotangent of dD/dt (sfh_rhs_vjp_augmented) or of D one step later
 * (sfh_step_vjp_augmented) and gD receives the cotangent of the current D.  sfh_rhs_vjp and
 * sfh_step_vjp return -3 on such a handle, these two on any other.  No counterpart in the
 * reference (AugmentedBoundaryIntegrator.cuh:32-36 is the forward system). */
int sfh_rhs_vjp_augmented(sfh_handle h, const double* gvx, const double* gvy, const double* gdphi,
                          const double* gdD, double* gx, double* gy, double* gphi, double* gD);
int sfh_step_vjp_augmented(sfh_handle h, double dt, const double* gx_in, const double* gy_in, const double* gphi_in,
                           const double* gD_in, double* gx_out, double* gy_out, double* gphi_out, double* gD_out);
/* Differentiable trajectories with device-resident checkpoints (no counterpart in the reference; the
 * forward map is sfh_step).  sfh_traj_forward advances `steps` RK4 steps of size dt from the current
 * state exactly as sfh_step(h, dt, ...) does, keeps in device memory the state before steps 0, every,
 * 2 every, ... with the solver's warm-start state there (ceil(steps / every) checkpoints of 32 T bytes,
 * 192 T bytes on GMRES and batched handles, T = sfh_total_nodes) and returns the id of
 * that record in *record; the handle's state afterwards is the final state.  sfh_traj_backward takes
 * the cotangent (gx, gy, gphi) of the state after the recorded steps to the cotangent of the state the
 * record started from: each segment, last first, is re-run from its checkpoint, keeping every step's
 * state, stage states and stage mu in device buffers (104 B per node per step), and swept back one
 * step VJP (sfh_step_vjp's reverse pass, on the kept stages) at a time, the cotangent staying in
 * device memory; the handle's state afterwards is the record's initial state.  A record can be swept
 * back any number of times until sfh_traj_release frees it (sfh_destroy frees every record).  Same
 * problem scope as sfh_step_vjp (-3 otherwise, before any step); -3 for an unknown record. */
int sfh_traj_forward(sfh_handle h, double dt, int steps, int checkpoint_every, int* record);
int sfh_traj_backward(sfh_handle h, int record, const double* gx_in, const double* gy_in, const double* gphi_in,
                      double* gx_out, double* gy_out, double* gphi_out);
int sfh_traj_release(sfh_handle h, int record);
/* Energies (EnergyContainer and VolumeFlux, Energies.cuh:136-204, computed by the reference on
 * every right-hand side, BaseBoundaryIntegrator.cuh:270-272).  While tracking is enabled every
 * right-hand side also writes its per-node energy terms (32 B per node; the step graph is
 * re-captured when tracking changes).  sfh_get_energies fills out[4 * s + c] for surface s with
 * c = 0 kinetic, 1 potential (gravitational for water, van der Waals for helium, times g),
 * 2 surface (kappa), 3 volume flux, of the last right-hand side evaluated: after sfh_step that is
 * the fourth stage of the last step, as the reference's getEnergy() after runStep; after sfh_rhs,
 * the current state.  Returns -3 when tracking is off. */
int sfh_track_energies(sfh_handle h, int enable);
int sfh_get_energies(sfh_handle h, double* out);
/* Dipole strengths of the last right-hand-side evaluation. */
int sfh_get_mu(sfh_handle h, double* mu);
int sfh_sync(sfh_handle h);
int sfh_get_stats(sfh_handle h, sfh_stats* out);
/* Device addresses of the state (complex128 Z, float64 phi) for zero-copy interop. */
int sfh_device_state(sfh_handle h, void** Z, void** phi);
/* Kernel timing.  While enabled, sfh_step / sfh_rhs launch without graph replay and bracket
 * every kernel launch with hipEvents on the solver's stream.  sfh_kernel_report sums, for one
 * kernel tag ("deriv", "assemble", "lu_step", "lu_backsolve", "mu_deriv", "velocity", and the GMRES
 * tags "gm_iter", "gm_close", "gm_close_mu"), the launches, milliseconds, algorithmic flops and algorithmic bytes recorded since
 * timing was enabled.  No counterpart in the reference (it has no timers). */
int sfh_kernel_timing(sfh_handle h, int enable);
int sfh_kernel_report(sfh_handle h, const char* tag, int* launches, double* ms, double* flops, double* bytes);

/* Right-hand side of ONE surface without any device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Computes what sfh_rhs returns for a fresh
 * handle of that surface created with SFH_SOLVE_BATCHED, SFH_FP64, tol and max_iter (zero initial
 * guess), from DEVICE arrays x, y, phi (n entries) to DEVICE arrays vx, vy, dphi and, if non-null,
 * mu, enqueued on `stream` (a hipStream_t; NULL for the default stream) without waiting for it.
 * Limits: n = 2^k or 3 * 2^k with 1 <= n <= 1024 (-2 otherwise).  Every intermediate (the n x n
 * FP64 matrix, derivatives, solve vectors, twiddles) lives in a static device array of
 * sfh_rhs_noalloc_scratch_bytes() bytes (about 8.5 MiB), one per GPU, held until the process
 * ends.  All calls on one GPU share it: calls that overlap on a GPU race.  Returns 0, -2, -3 or -1
 * (launch error). */
int sfh_rhs_noalloc(const sfh_props* props, int n, const double* x, const double* y, const double* phi, double* vx,
                    double* vy, double* dphi, double* mu, double tol, int max_iter, void* stream);
size_t sfh_rhs_noalloc_scratch_bytes(void);

/* Message of the last failed call on this thread. */
const char* sfh_last_error(void);

/* Test hook: solves A x = b (n x n row-major host arrays) with the device LU path of the
 * stepper (interchange-free blocked LU, GEPP kernel when a multiplier exceeds 1, reported in
 * *interchanged).  No counterpart in the reference; it exposes MatrixSolver::solve semantics
 * (MatrixSolver.cuh:113-126) on arbitrary matrices. */
int sfh_debug_lu_solve(int n, const double* A, const double* b, double* x, int* interchanged);

#ifdef __cplusplus
}
#endif

#endif /* CUSUPERHELIUM_H */
