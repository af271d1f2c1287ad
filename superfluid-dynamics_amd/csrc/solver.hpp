// solver.hpp -- host-side ensemble solver: device buffers, RHS pipeline and RK4 driver.
//
// Mirrors the reference object graph BaseBoundaryIntegralCalculator<N,B> + BoundaryProblem +
// AutonomousRungeKuttaStepper (CuSuperHelium/BaseBoundaryIntegrator.cuh:10-85,
// AutonomousRungeKuttaStepper.cuh:24-89) with run-time sizes instead of template parameters:
// one Solver owns S surfaces of arbitrary sizes, allocates everything once, and advances the
// state in HBM without host round trips.
#pragma once
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "sfh_kernels.hpp"

namespace sfh {

#define SFH_HIP(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            throw ::sfh::HipError(std::string(#expr) + ": " + hipGetErrorString(e_));          \
    } while (0)

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct ArgError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
// A right-hand side came out NaN or infinite (a NaN state given on the device, coincident nodes,
// a blown-up trajectory); the handle's state is then invalid until the next set_state.
struct NonFiniteError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

enum Precision { kFP64 = 0, kMixed = 1 };
enum SolveKind { kSolveLU = 0, kSolveGMRES = 1, kSolveBatched = 2 };

struct Options {
    int precision = kFP64;
    int solver = kSolveLU;
    int use_graph = 1;
    int max_iter = 32;       // GMRES Arnoldi steps per solve (at most kGmMaxIter) / batched GMRES cap
    int restart = 0;         // unused: the multi-workgroup GMRES does not restart, the batched one
                             // restarts every 8 vectors
    int force_trim = 0;      // test hook: trimmed GMRES sequence length instead of the calibrated one
    int pair_rows = 0;       // rows per workgroup of the pair kernels: 1, 2 or 4 (0: 4)
    double tol = 1e-12;      // GMRES relative residual / refinement tolerance
};

// Most Arnoldi steps of one multi-workgroup GMRES solve (its launch sequence has no restart).
constexpr int kGmMaxIter = 32;

struct Stats {
    long long rhs_evals = 0;
    long long steps = 0;
    int fallback_solves = 0;   // interchange-needing systems re-solved by GEPP
    int spin_timeouts = 0;
    int last_iterations = 0;   // iterative solvers: iterations of the last solve (surface 0)
    int max_iterations = 0;    // iterative solvers: most iterations any surface used last solve
    int unconverged = 0;       // iterative solves that hit the iteration cap
    int refinements = 0;       // mixed precision: refinement sweeps of the last solve (surface 0)
    int redone_steps = 0;      // GMRES: steps discarded and run again with a longer launch sequence
    int sequence_length = 0;   // GMRES: launches per solve of the trimmed sequence (0 while calibrating)
};

bool supported_n(int n);
void debug_lu_solve(int n, const double* A, const double* b, double* x, int* interchanged);

// Autotuner (autotune.cpp): the fastest (solver, precision, pair-kernel rows) for one surface of n
// nodes of this physics on the current GPU, measured; tuned() returns the cached result or tunes.
struct Tuning {
    int solver = kSolveLU;
    int precision = kFP64;
    int pair_rows = kPairRows;
    int from_cache = 0;
    double ms_per_step = 0.0;
};
Tuning autotune(const Phys& ph, int n, bool allow_mixed, int steps);
Tuning tuned(const Phys& ph, int n, bool allow_mixed);

// The +eps perturbed states of calculatePerturbedStates256 (deriv.hip): out holds 6 n^2 complex.
void perturbed_states(const double* x, const double* y, const double* phi, int n, double eps, double* out_host);

// Central-difference Jacobian of [vx, vy, dphi] with respect to the real state [x, y, phi] of one
// surface of n <= 2048 nodes: jac[k 3n + r] = d f_r / d s_k (implicit.cpp).
void fd_jacobian(const Phys& ph, int n, const double* state, double eps, double* jac);

// Two-stage Gauss-Legendre integration of one surface (n <= 1024) from t0 to t1 (implicit.cpp).
struct Gl2Options {
    double tol = 1e-10;          // Newton: |R| <= tol (1 + |k|^2)
    int max_newton = 20;
    int max_halves = 6;
    bool simplified_fallback = false;
    double armijo_c = 1e-4, backtrack = 0.5, min_alpha = 1e-6;
    double eps = 1e-6;           // Jacobian perturbation
};
int gl2_integrate(const Phys& ph, int n, double* state, double t0, double t1, double h, const Gl2Options& o,
                  std::vector<double>* traj, std::vector<double>* times);
// Right-hand side of one surface (n <= 1024) from device arrays to device arrays on `stream`, every
// intermediate in a static device array (noalloc.hip): no device memory is allocated.
void rhs_noalloc(const Phys& ph, int n, const double* x, const double* y, const double* phi, double* vx,
                 double* vy, double* dphi, double* mu, double tol, int max_iter, hipStream_t stream);
size_t noalloc_scratch_bytes();

class Solver {
  public:
    Solver(const Phys& ph, const std::vector<int>& n, const Options& opt);
    ~Solver();
    Solver(const Solver&) = delete;
    Solver& operator=(const Solver&) = delete;

    int surfaces() const { return (int)n_.size(); }
    int total_nodes() const { return T_; }
    const std::vector<int>& sizes() const { return n_; }
    hipStream_t stream() const { return stream_; }

    void set_state(const double* x, const double* y, const double* phi);
    // Device-resident state in the reference's layout, complex128 [Z(T), Phi(T) + 0i]
    // (BaseBoundaryIntegrator.cuh:141-146).  rhs_device is AutonomousProblem::run (AutonomousProblem.h:25):
    // rhs = [dZ/dt (T), dPhi/dt + 0i (T)] of `state`, the handle's own state left untouched; all
    // three are enqueued on the handle's stream and complete before the call returns.
    void set_state_device(const double2* state);
    // Optical drive (HeliumDrivenAutonomousProblem, HeliumDrivenAutonomousProblem.cuh:19-25) and,
    // with ph.augmented, the delayed intensity D as a third state field (AugmentedBoundaryIntegrator,
    // AugmentedBoundaryIntegrator.cuh:32-36).  set_drive takes the drive fields of ph (the rest of
    // ph is kept), zeroes D and re-captures the step graph.  D is zeroed by set_state too.
    void set_drive(const Phys& drive);
    void set_delayed(const double* D);
    void get_delayed(double* D);
    void get_delayed_rhs(double* dD);   // dD/dt of the last standalone right-hand side (rhs())
    void get_state_device(double2* state);
    void rhs_device(const double2* state, double2* rhs);
    void get_state(double* x, double* y, double* phi);
    // Right-hand side of the current state: vx + i vy = dZ/dt, dphi = dPhi/dt (host arrays).
    void rhs(double* vx, double* vy, double* dphi, double* mu = nullptr, double2* zp = nullptr,
             double2* zpp = nullptr, double2* qup = nullptr);
    void step(double dt, int nsteps);
    // nsteps steps recording the state after every `every`-th one (TrajectoryLogger::logTrajectory,
    // TrajectoryLogger.cuh:65-73): frame f = [x(T), y(T), phi(T)] into states[f * 3 T], the handle's
    // time after that step into times[f].  Frames go through a device ring and are copied to host
    // memory on a second stream while the stepping continues.
    void step_record(double dt, int nsteps, int every, double* states, double* times);
    double time() const { return time_; }
    void get_mu(double* mu);
    void sync();
    Stats stats();
    void clear_solver_state();
    // Sets the handle's time (and, with the timed drive, the start of its delayed-intensity clock).
    void set_time(double t);
    // The form the handle runs: solver, precision and pair-kernel rows after defaults are applied.
    int solver() const { return opt_.solver; }
    int precision() const { return opt_.precision; }
    int pair_rows() const { return rows_; }

    // Raw device buffers (state y0) for zero-copy interop.
    double2* device_Z() { return Z0_; }
    double* device_phi() { return P0_; }

    // Reverse mode (water, the helium film, the autonomous drive).  rhs_vjp: cotangents of (vx, vy, dphi) at the
    // current state -> cotangents of (x, y, phi).  step_vjp: cotangent of the state after one RK4
    // step of size dt taken from the current state -> cotangent of the current state (the state is
    // left unchanged).  Host arrays of total_nodes() entries.
    // With the augmented drive the state has a third field D: gdD / gD_in is its cotangent in and
    // gD / gD_out out (both required then, and refused otherwise).
    void rhs_vjp(const double* gvx, const double* gvy, const double* gdphi, double* gx, double* gy, double* gphi,
                 const double* gdD = nullptr, double* gD = nullptr);
    void step_vjp(double dt, const double* gx_in, const double* gy_in, const double* gp_in, double* gx_out,
                  double* gy_out, double* gp_out, const double* gD_in = nullptr, double* gD_out = nullptr);

    // Differentiable trajectories with device-resident checkpoints.  traj_forward advances `steps`
    // RK4 steps of size dt from the current state exactly as step() does, keeping in device memory the
    // state before steps 0, every, 2 every, ...; it returns the id of that record.  traj_backward takes
    // the cotangent of the state after the recorded steps to the cotangent of the state it started
    // from: each segment, last first, is re-run from its checkpoint, keeping every step's state, stage
    // states and stage mu in device buffers, then swept back one RK4 step at a time from those (no
    // solve besides the transposed ones), the cotangent kept on the device.  The handle's state
    // afterwards is the record's initial state.  Host arrays of total_nodes() entries.
    int traj_forward(double dt, int steps, int every);
    void traj_backward(int record, const double* gx_in, const double* gy_in, const double* gp_in, double* gx_out,
                       double* gy_out, double* gp_out);
    void traj_release(int record);

    // Energies (EnergyContainer + VolumeFlux, Energies.cuh:136-204): while tracking is on, every
    // right-hand side also writes its per-node energy terms; energies() returns, per surface,
    // {kinetic, potential, surface, volume flux} of the last right-hand side evaluated.
    void track_energies(bool on);
    void energies(double* out);

    // Standalone derivative of a complex field (FftDerivative::exec semantics).
    void derivative(const double2* in_host, double2* out_host, bool second, double scale);

    // Kernel timing: when enabled, steps run without graph replay and every launch is bracketed
    // by hipEvents on the solver's stream; report() sums them per kernel tag.
    void set_kernel_timing(bool on);
    struct KernelReport {
        int launches = 0;
        double ms = 0.0;
        double flops = 0.0;
        double bytes = 0.0;
    };
    KernelReport kernel_report(const std::string& tag);

  private:
    void enqueue_rhs(const double2* Z, const double* P, const StageCtl& ctl);
    bool enqueue_solve(const double2* Z, int stage, bool g1);
    // the stage states Y2..Y4 and the four stage mu of one RK4 step, copied out as the step runs
    struct StageCapture {
        double2* Y[3];
        double* YP[3];
        double* mu[4];
    };
    void enqueue_step(double dt, const StageCapture* cap = nullptr);
    void capture(double dt, int len);
    void run_steps(double dt, int len, int count);   // count steps with GMRES sequences of len launches
    void calibrate();               // bookkeeping after a full-length step while the history fills
    void save_checkpoint();         // state, last solution and warm-start history, device to device
    void restore_checkpoint();
    void check_status();
    template <class F>
    void launch(const char* tag, double flops, double bytes, F&& f, int iter = -1);
    // A kernel launch on stream_.  Inside a timed launch() it records the kernel's own start (the
    // group's first kernel) and end into that launch's event pair (hipExtLaunchKernelGGL), as a
    // profiler times a kernel.
    template <typename K, typename... Args>
    void go(K kernel, dim3 grid, dim3 block, unsigned shm, Args... args);
    hipEvent_t ev_a_ = nullptr, ev_b_ = nullptr;

    struct Timed {
        const char* tag;
        hipEvent_t a, b;
        double flops, bytes;
        int iter;      // GMRES iteration of the launch (-1 otherwise)
        bool active;   // false for GMRES launches that returned at once after convergence
    };
    bool ktime_ = false;
    std::vector<Timed> timed_;
    std::vector<hipEvent_t> event_pool_;
    size_t event_next_ = 0;

    Phys ph_;
    Options opt_;
    int rows_ = kPairRows;   // rows per workgroup of the pair kernels
    std::vector<int> n_;
    std::vector<SurfDesc> desc_;
    int S_ = 0, T_ = 0, maxN_ = 0;
    long long MT_ = 0;
    bool uniform_ = true;
    hipStream_t stream_ = nullptr;

    SurfDesc* d_desc_ = nullptr;
    double2* d_tw_ = nullptr;
    double* d_circ_ = nullptr;   // circulant derivative kernels (multi-workgroup path)
    double2 *Z0_ = nullptr, *Za_ = nullptr, *Zb_ = nullptr;
    double *P0_ = nullptr, *Pa_ = nullptr, *Pb_ = nullptr;
    double2* kZ_[3] = {nullptr, nullptr, nullptr};
    double* kP_[3] = {nullptr, nullptr, nullptr};
    double2 *Zp_ = nullptr, *Zpp_ = nullptr, *mup_ = nullptr, *qup_ = nullptr;
    double *b_ = nullptr, *work_ = nullptr, *y_ = nullptr, *mu_ = nullptr;
    double* M_ = nullptr;
    NodeTrig* trig_ = nullptr;
    double* uinv_ = nullptr;
    int* flags_ = nullptr;
    int* status_ = nullptr;
    int* piv_ = nullptr;
    int* counters_ = nullptr;   // [0] fallback solves
    std::vector<long long> uoff_, foff_;
    // GMRES workspace
    int gm_m_ = 0, gm_nwg_ = 0;
    double *gmV_ = nullptr, *gmw_ = nullptr, *gmpart_ = nullptr, *gmsmall_ = nullptr;
    double* gmx0_ = nullptr;     // initial guess of the solve in flight (the solution goes to mu_)
    float* M32_ = nullptr;       // FP32 matrices of the mixed-precision GMRES form
    int* gmctl_ = nullptr;
    long long gm_part_stride_ = 0, gm_small_stride_ = 0;
    bool enqueue_gmres(const GmArgs& a, bool g1 = false, bool force64 = false);
    void alloc_gmres(int m);
    GmArgs gm_args(int stage) const;
    double* gmhist_ = nullptr;   // 4 x kHistDepth x total: per RK stage, mu of the last steps
    int* gmhvalid_ = nullptr;
    // persistent per-surface stage kernel (kSolveBatched)
    int* order_ = nullptr;    // workgroup -> surface, by decreasing N
    int* bstats_ = nullptr;   // 4 ints per surface
    double* bhist_ = nullptr;  // per RK stage, each surface's mu of the last kHistDepth steps
    int* bhvalid_ = nullptr;
    void enqueue_batched(const double2* Z, const double* P, const StageCtl& ctl);
    // adjoint workspace (allocated on first use)
    void ensure_vjp();
    // stage: the RK stage (0..3) of the reverse step, which selects the warm-start ring of the
    // transposed solve inside a trajectory sweep; -1 outside one
    void enqueue_vjp(const double2* Z, const double* P, const double* mu, const double2* gAZ, const double* gAP,
                     double ca, const double2* gBZ, const double* gBP, double cb, double2* outZ, double* outP,
                     bool accumulate, int stage = -1, const double* gAD = nullptr, const double* gBD = nullptr,
                     double* outD = nullptr);
    void upload_cotangent(const double* gx, const double* gy, const double* gp, double2* Z, double* P);
    void download_cotangent(const double2* Z, const double* P, double* gx, double* gy, double* gp);
    bool vjp_ready_ = false;
    double2 *vY_[3] = {nullptr, nullptr, nullptr}, *vK_[3] = {nullptr, nullptr, nullptr};
    double *vYP_[3] = {nullptr, nullptr, nullptr}, *vKP_[3] = {nullptr, nullptr, nullptr};
    double* vMu_[4] = {nullptr, nullptr, nullptr, nullptr};
    double2 *gA_ = nullptr, *gR_[2] = {nullptr, nullptr}, *gAcc_ = nullptr;
    double *gAP_ = nullptr, *gRP_[2] = {nullptr, nullptr}, *gAccP_ = nullptr;
    AdjNode* adj_ = nullptr;
    int* vstats_ = nullptr;
    // multi-workgroup reverse mode (LU and GMRES handles): M^T and the per-node vectors of adjoint_mwg.hip
    double* Mt_ = nullptr;
    double2 *agwc_ = nullptr, *agmup_ = nullptr, *ad1_ = nullptr, *agam_ = nullptr, *agamI_ = nullptr;
    double2 *agzp_ = nullptr, *agzpp_ = nullptr, *agz_ = nullptr;
    double *agmu_ = nullptr, *agb_ = nullptr;
    double* aquirk_ = nullptr;   // the V1[1] quirk terms of the reverse pass (3 x T_)
    // the augmented drive's D in the reverse pass: input cotangent, accumulator, stage cotangents,
    // recomputed stage values (allocated by ensure_vjp_d)
    double *gAD_ = nullptr, *gAccD_ = nullptr, *gRD_[2] = {nullptr, nullptr}, *vYD_[3] = {nullptr, nullptr, nullptr};
    void ensure_vjp_d();
    void enqueue_vjp_mwg(const double2* Z, const double* P, const double* mu, const double2* gAZ, const double* gAP,
                         double ca, const double2* gBZ, const double* gBP, double cb, double2* outZ, double* outP,
                         bool accumulate, int stage);
    // warm starts of the transposed solves over a trajectory's reverse steps (LU and GMRES handles):
    // per RK stage a ring of the last kHistDepth solutions, extrapolated like the forward warm start
    double* adjRing_ = nullptr;      // 4 stages x kHistDepth x T
    double* adjX_ = nullptr;         // the extrapolated guess
    double* adjSave_ = nullptr;      // the rings at the start of a segment (for a redo)
    double2* adjSaveZ_ = nullptr;    // the cotangent at the start of a segment
    double* adjSaveP_ = nullptr;
    int adj_head_[4] = {0, 0, 0, 0}, adj_hv_[4] = {0, 0, 0, 0};
    bool adj_warm_ = false;          // inside a trajectory sweep
    int adj_len_ = 0;                // launches per warm transposed solve (0: the full sequence)
    bool adj_calib_ = false;         // this reverse step measures the Arnoldi steps of its solves
    int adj_calib_max_ = 0;
    // one RK4 step back from the state (Z, P): cotangent gA_/gAP_ of the state after it -> gAcc_/gAccP_.
    // The stage states and mu come from `stored` when given, else they are recomputed (solves from
    // zero guesses with the full launch sequence).
    void enqueue_step_vjp(double dt, const double2* Z, const double* P, const StageCapture* stored);
    // the state from device buffers, with the resets of set_state (no host copy, no wait)
    void load_state(const double2* Z, const double* P);
    struct TrajRecord {
        double dt = 0.0;
        int steps = 0, every = 1;
        double2* Z = nullptr;   // checkpoint c at Z + c T
        double* P = nullptr;
        // the solver's warm-start state at each checkpoint: the last mu, and (GMRES, batched) the
        // stage histories and their valid counts, so a segment's re-run starts where the forward pass was
        double* Mu = nullptr;   // c T
        double* H = nullptr;    // c * hist_len
        int* HV = nullptr;      // c * hv_len
        size_t hist_len = 0, hv_len = 0;
        std::vector<int> hist_steps;   // hist_steps_ at each checkpoint
    };
    void free_record(TrajRecord& r);
    std::map<int, TrajRecord> traj_;
    int traj_next_ = 1;
    double2* segZ_ = nullptr;   // the states of one segment
    double* segP_ = nullptr;
    double2* segY_ = nullptr;   // their stage states (3 per step) and stage mu (4 per step)
    double* segYP_ = nullptr;
    double* segMu_ = nullptr;
    int seg_cap_ = 0;
    StageCapture seg_capture(int k) const;

    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    double graph_dt_ = 0.0;
    int graph_len_ = 0;
    // GMRES launch sequence per solve.  The step graph holds a fixed number of Arnoldi launches per
    // solve (launches after convergence return at once but still cost a dispatch each).  While the
    // warm-start history fills (the first kHistDepth steps after set_state) and for one step with
    // the full history, every solve gets the full cap gm_m_; afterwards the sequence is trimmed to
    // what the solves of that last step needed (one launch per Arnoldi step, one to finish, a spare
    // step when they needed 3 or more), and doubled (up to the cap) whenever a check finds a solve
    // that ran out of launches.  seq_len_ is the length the kernels are being enqueued with.
    // A trimmed sequence is run in chunks of steps.  Before a chunk the state, the last solution and
    // the warm-start history are copied aside; when a solve of the chunk ran out of launches the
    // copy is restored, the length doubled (up to the cap) and the chunk run again, so every step
    // that is kept had all of its solves converge (or ran with the full cap).
    int seq_len_ = 0;
    int trim_ = 0;          // 0: not determined yet
    int hist_steps_ = 0;    // steps since set_state (saturates at kCalibSteps)
    int chunk_ = 2;         // steps of the next trimmed chunk (doubles up to kCheckSteps)
    int unc_raw_ = 0;       // unconverged count of the control blocks (discarded chunks included)
    int unc_discard_ = 0;   // ... of solves in chunks that were run again
    int fill_iters_ = 0;    // most Arnoldi steps of a solve with the full history, any surface
    double2* ckZ_ = nullptr;
    double *ckP_ = nullptr, *ckMu_ = nullptr, *ckHist_ = nullptr;
    int* ckHv_ = nullptr;
    Stats stats_;
    double* energy_ = nullptr;   // 4 doubles per node while energies are tracked
    bool track_energy_ = false;
    double time_ = 0.0;          // advanced by dt per step, 0 after set_state
    // timed drive: the clock of the stage times (t_n of the next step) and the time of the last
    // saved delayed intensity (setStartingTime / save_value, DelayedIntensityTerm.cuh:31-37,66-69),
    // and their values at the chunk checkpoint
    double drive_t_ = 0.0, drive_prev_ = 0.0, ck_drive_t_ = 0.0, ck_drive_prev_ = 0.0;
    void timed_coefficients(StageCtl& c, double ts, int save) const;
    // augmented drive: delayed intensity of y0, the stage states, the slopes, the checkpoint copy
    double *D0_ = nullptr, *Da_ = nullptr, *Db_ = nullptr, *ckD_ = nullptr;
    double* kD_[3] = {nullptr, nullptr, nullptr};
    void set_delayed_ctl(StageCtl& c, const double* Din, double* Dout) const;
    // trajectory recording (allocated on first use): device ring of frames, pinned host staging,
    // copy stream and one event pair per slot
    int rec_slots_ = 0;
    size_t rec_frame_ = 0;
    double* rec_ring_ = nullptr;
    double* rec_host_ = nullptr;
    hipStream_t rec_stream_ = nullptr;
    std::vector<hipEvent_t> rec_packed_, rec_copied_;
};

}  // namespace sfh
