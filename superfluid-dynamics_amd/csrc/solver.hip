// solver.hip -- host orchestration of the gfx950 Roberts RK4 stepper.
//
// One RHS evaluation (BaseBoundaryIntegralCalculator::runTimeStep,
// CuSuperHelium/BaseBoundaryIntegrator.cuh:137-306) is enqueued as
//   zphi_circ_kernel      Z', Z'', Re Phi'            (circulant products, 8 rows per workgroup)
//   assemble64_kernel     M                           (8 rows per workgroup)
//   LU solve              mu                          (panel / MFMA update / back substitution)
//   GEPP fallback         no-op unless a system needed row interchanges
//     or GMRES            mu                          (one kernel per Arnoldi step; the last
//                                                      launch also forms mu', gm_close_mu_kernel)
//   mu_circ_kernel        mu'                         (LU, and GMRES with a one-launch sequence)
//   velocity_stage_kernel q, dPhi/dt and the RK4 stage update of every node
//     or velocity_gm1_kernel                          (FP64 GMRES sequence of two launches: the
//                                                      product, then this kernel, which finishes
//                                                      the step and forms mu and mu' as well)
// and one RK4 step (AutonomousRungeKuttaStepper.cuh:123-361) is four of them.  The step is
// captured once into a hipGraph and replayed, so a run of many steps costs one graph launch
// per step and the state never leaves HBM.
#include <algorithm>
#include <cmath>
#include <cstring>

#include <hip/hip_ext.h>

#include "sfh_gmres.hpp"
#include "solver.hpp"

namespace sfh {

// Launches a kernel templated on the rows per workgroup R of the pair kernels (1, 2 or 4) with the
// run-time choice rows_: the statement sees R as a constant.
#define SFH_WITH_ROWS(rows_, ...)              \
    switch (rows_) {                           \
    case 1: { constexpr int R = 1; __VA_ARGS__; } break; \
    case 2: { constexpr int R = 2; __VA_ARGS__; } break; \
    default: { constexpr int R = 4; __VA_ARGS__; } break; \
    }

// Steps that run every solve with the full launch cap: kHistDepth while the warm-start history
// fills, then one with the full history that sets the trimmed length.
constexpr int kCalibSteps = kHistDepth + 1;
// Longest chunk of trimmed steps between two checks.
constexpr int kCheckSteps = 32;

bool supported_n(int n) {
    if (n < 1 || n > 8192) return false;
    int m = n;
    if (m % 3 == 0) m /= 3;
    return (m & (m - 1)) == 0;
}

// Circulant kernels of the spectral derivatives for n nodes, appended to out as Re c1, Im c1, c2
// (n doubles each): c_q = sum_i m_i exp(+2 pi i i q / n) over the multiplier table of
// utilities.cuh:119-147 (first derivative: i m_i with m_i = i / n below n/2, pi i / n at n/2,
// 0 at n/2 + 1, (i - n) / n above) and :185-199 (second: -i^2 / n up to n/2, -(i - n)^2 / n
// above), summed in long double with exactly reduced angles.
static void circulant_tables(int n, std::vector<double>& out) {
    const long double pi = 3.141592653589793238462643383279502884L;
    std::vector<long double> cs(n), sn(n), m1(n), m2(n);
    for (int t = 0; t < n; ++t) {
        const long double a = 2.0L * pi * (long double)t / (long double)n;
        cs[t] = cosl(a);
        sn[t] = sinl(a);
    }
    for (int i = 0; i < n; ++i) {
        const long double ln = (long double)n;
        if (i < n / 2) m1[i] = (long double)i / ln;
        else if (i == n / 2) m1[i] = pi * (long double)i / ln;
        else if (i == n / 2 + 1) m1[i] = 0.0L;
        else m1[i] = (long double)(i - n) / ln;
        const long double f = (i <= n / 2) ? (long double)i : (long double)(i - n);
        m2[i] = -f * f / ln;
    }
    const size_t base = out.size();
    out.resize(base + 3 * (size_t)n);
    for (int q = 0; q < n; ++q) {
        long double re = 0.0L, im = 0.0L, sec = 0.0L;
        long long t = 0;   // i q mod n
        for (int i = 0; i < n; ++i) {
            re -= m1[i] * sn[t];   // i m (cos + i sin) = -m sin + i m cos
            im += m1[i] * cs[t];
            sec += m2[i] * cs[t];  // m2 is symmetric: the sine terms cancel
            t += q;
            if (t >= n) t -= n;
        }
        out[base + q] = (double)re;
        out[base + n + q] = (double)im;
        out[base + 2 * (size_t)n + q] = (double)sec;
    }
}

static SurfDesc make_desc(int n) {
    SurfDesc d{};
    d.n = n;
    int m = n;
    if (m % 3 == 0) {
        d.r3 = 1;
        m /= 3;
    }
    int l = 0;
    while ((1 << l) < m) ++l;
    d.log2m = l;
    return d;
}

template <typename T>
static T* dalloc(size_t count) {
    T* p = nullptr;
    if (count == 0) count = 1;
    SFH_HIP(hipMalloc(&p, count * sizeof(T)));
    return p;
}

Solver::Solver(const Phys& ph, const std::vector<int>& n, const Options& opt) : ph_(ph), opt_(opt), n_(n) {
    if (n.empty()) throw ArgError("no surfaces");
    S_ = (int)n.size();
    for (int v : n)
        if (!supported_n(v)) throw ArgError("unsupported node count " + std::to_string(v));
    if (opt_.solver != kSolveLU && opt_.solver != kSolveGMRES && opt_.solver != kSolveBatched)
        throw ArgError("unknown solver");
    if (opt_.precision != kFP64 && opt_.precision != kMixed) throw ArgError("unknown precision");
    if (opt_.precision == kMixed && opt_.solver == kSolveLU)
        throw ArgError("the mixed-precision form runs with the GMRES or the batched solver");
    if (opt_.solver == kSolveGMRES) {
        for (int v : n)
            if (v > kGmMaxN) throw ArgError("the GMRES solver takes surfaces of at most 4096 nodes");
        if (opt_.max_iter < 1 || opt_.max_iter > kGmMaxIter)
            throw ArgError("the GMRES solver runs 1 to " + std::to_string(kGmMaxIter) +
                           " Arnoldi steps per solve (max_iter), without restart");
    }
    if (opt_.force_trim < 0) throw ArgError("negative force_trim");
    if (opt_.pair_rows != 0 && opt_.pair_rows != 1 && opt_.pair_rows != 2 && opt_.pair_rows != 4)
        throw ArgError("pair_rows must be 1, 2 or 4");
    rows_ = opt_.pair_rows ? opt_.pair_rows : kPairRows;
    if (opt_.solver == kSolveBatched)
        for (int v : n)
            if (v > kBatchMaxN) throw ArgError("the batched solver takes surfaces of at most 1024 nodes");
    SFH_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));

    // descriptors, twiddle tables and circulant derivative kernels (one table each per distinct
    // n, computed in long double)
    std::vector<double2> tw;
    std::vector<double> circ;
    std::vector<std::pair<int, int>> tw_of, circ_of;
    long long moff = 0;
    int voff = 0;
    for (int s = 0; s < S_; ++s) {
        SurfDesc d = make_desc(n[s]);
        d.voff = voff;
        d.moff = moff;
        int off = -1;
        for (auto& p : tw_of)
            if (p.first == n[s]) off = p.second;
        if (off < 0) {
            off = (int)tw.size();
            tw_of.push_back({n[s], off});
            for (int t = 0; t < n[s]; ++t) {
                const long double ang =
                    2.0L * 3.141592653589793238462643383279502884L * (long double)t / (long double)n[s];
                tw.push_back(make_double2((double)cosl(ang), (double)(-sinl(ang))));
            }
        }
        d.twoff = off;
        if (opt_.solver != kSolveBatched) {
            int coff = -1;
            for (auto& p : circ_of)
                if (p.first == n[s]) coff = p.second;
            if (coff < 0) {
                coff = (int)circ.size();
                circ_of.push_back({n[s], coff});
                circulant_tables(n[s], circ);
            }
            d.coff = coff;
        }
        desc_.push_back(d);
        voff += n[s];
        // the batched kernel keeps each surface's matrix in 64 x 64 tiles
        moff += opt_.solver == kSolveBatched ? tiled_matrix_elems(n[s]) : (long long)n[s] * n[s];
        maxN_ = std::max(maxN_, n[s]);
        if (n[s] != n[0]) uniform_ = false;
    }
    T_ = voff;
    MT_ = moff;

    d_desc_ = dalloc<SurfDesc>(S_);
    SFH_HIP(hipMemcpy(d_desc_, desc_.data(), S_ * sizeof(SurfDesc), hipMemcpyHostToDevice));
    d_tw_ = dalloc<double2>(tw.size());
    SFH_HIP(hipMemcpy(d_tw_, tw.data(), tw.size() * sizeof(double2), hipMemcpyHostToDevice));
    if (!circ.empty()) {
        d_circ_ = dalloc<double>(circ.size());
        SFH_HIP(hipMemcpy(d_circ_, circ.data(), circ.size() * sizeof(double), hipMemcpyHostToDevice));
    }

    Z0_ = dalloc<double2>(T_);
    Za_ = dalloc<double2>(T_);
    Zb_ = dalloc<double2>(T_);
    P0_ = dalloc<double>(T_);
    Pa_ = dalloc<double>(T_);
    Pb_ = dalloc<double>(T_);
    for (int i = 0; i < 3; ++i) {
        kZ_[i] = dalloc<double2>(T_);
        kP_[i] = dalloc<double>(T_);
    }
    Zp_ = dalloc<double2>(T_);
    Zpp_ = dalloc<double2>(T_);
    mup_ = dalloc<double2>(T_);
    qup_ = dalloc<double2>(T_);
    b_ = dalloc<double>(T_);
    work_ = dalloc<double>(T_);
    y_ = dalloc<double>(T_);
    mu_ = dalloc<double>(T_);
    M_ = dalloc<double>(MT_);
    trig_ = dalloc<NodeTrig>(T_);
    long long utot = 0, ftot = 0;
    for (int s = 0; s < S_; ++s) {
        const int P = (n[s] + kNB - 1) / kNB;
        uoff_.push_back(utot);
        foff_.push_back(ftot);
        utot += (long long)P * kNB * kNB;
        ftot += P;
    }
    uinv_ = dalloc<double>(utot);
    flags_ = dalloc<int>(ftot);
    status_ = dalloc<int>(2 * S_);
    piv_ = dalloc<int>(T_);
    counters_ = dalloc<int>(4);
    SFH_HIP(hipMemset(status_, 0, 2 * S_ * sizeof(int)));
    SFH_HIP(hipMemset(counters_, 0, 4 * sizeof(int)));
    SFH_HIP(hipMemset(Z0_, 0, T_ * sizeof(double2)));
    SFH_HIP(hipMemset(P0_, 0, T_ * sizeof(double)));
    SFH_HIP(hipMemset(mu_, 0, T_ * sizeof(double)));
    SFH_HIP(hipMemset(mup_, 0, T_ * sizeof(double2)));
    if (opt_.solver == kSolveGMRES) alloc_gmres(opt_.max_iter);
    if (opt_.solver == kSolveBatched) {
        std::vector<int> ord(S_);
        for (int s = 0; s < S_; ++s) ord[s] = s;
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return n_[a] > n_[b]; });
        order_ = dalloc<int>(S_);
        SFH_HIP(hipMemcpy(order_, ord.data(), S_ * sizeof(int), hipMemcpyHostToDevice));
        bstats_ = dalloc<int>(4 * (size_t)S_);
        SFH_HIP(hipMemset(bstats_, 0, 4 * S_ * sizeof(int)));
        bhist_ = dalloc<double>(4 * kHistDepth * (size_t)T_);
        bhvalid_ = dalloc<int>(4 * (size_t)S_);
        SFH_HIP(hipMemset(bhvalid_, 0, 4 * S_ * sizeof(int)));
    }
    SFH_HIP(hipDeviceSynchronize());
}

// GMRES buffers for a Krylov dimension of m (the GMRES solver, and the transposed solves of the
// multi-workgroup adjoint of the other solvers)
void Solver::alloc_gmres(int m) {
    if (gmV_) return;
    gm_m_ = m;   // at most kMaxKrylov basis vectors (gmres.hip)
    gm_nwg_ = (maxN_ + 3) / 4;   // kRowsPerWG in gmres.hip
    gm_part_stride_ = 2LL * gm_nwg_ * (gm_m_ + 1);
    gm_small_stride_ = 4LL * (gm_m_ + 1) + 2LL * gm_m_ + (long long)gm_m_ * (gm_m_ + 1);
    gmV_ = dalloc<double>((size_t)(gm_m_ + 1) * T_);
    gmx0_ = dalloc<double>(T_);
    if (opt_.precision == kMixed && opt_.solver == kSolveGMRES) M32_ = dalloc<float>(MT_);
    SFH_HIP(hipMemset(gmx0_, 0, T_ * sizeof(double)));
    gmw_ = dalloc<double>(2 * (size_t)T_);
    gmpart_ = dalloc<double>(gm_part_stride_ * S_);
    gmsmall_ = dalloc<double>(gm_small_stride_ * S_);
    gmctl_ = dalloc<int>((size_t)kGmCtlInts * S_);
    SFH_HIP(hipMemset(gmctl_, 0, sizeof(int) * kGmCtlInts * S_));
    SFH_HIP(hipMemset(gmsmall_, 0, sizeof(double) * gm_small_stride_ * S_));
    gmhist_ = dalloc<double>(4 * kHistDepth * (size_t)T_);
    gmhvalid_ = dalloc<int>(4);
    SFH_HIP(hipMemset(gmhvalid_, 0, 4 * sizeof(int)));
    ckZ_ = dalloc<double2>(T_);
    ckP_ = dalloc<double>(T_);
    ckMu_ = dalloc<double>(T_);
    ckHist_ = dalloc<double>(4 * kHistDepth * (size_t)T_);
    ckHv_ = dalloc<int>(4);
}

Solver::~Solver() {
    // replays of the step graph may still be in flight: the executable graph is destroyed only
    // once the stream has drained
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    void* bufs[] = {d_desc_, d_tw_, Z0_, Za_, Zb_, P0_, Pa_, Pb_, kZ_[0], kZ_[1], kZ_[2], kP_[0],
                    kP_[1], kP_[2], Zp_, Zpp_, mup_, qup_, b_, work_, y_, mu_, M_, uinv_, flags_,
                    status_, piv_, counters_, gmV_, gmw_, gmpart_, gmsmall_, gmctl_, trig_, order_, bstats_,
                    gmhist_, gmhvalid_, bhist_, bhvalid_, d_circ_, gmx0_, M32_, ckZ_, ckP_, ckMu_, ckHist_,
                    ckHv_, energy_, D0_, Da_, Db_, ckD_, kD_[0], kD_[1], kD_[2]};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    void* vbufs[] = {vY_[0], vY_[1], vY_[2], vK_[0], vK_[1], vK_[2], vYP_[0], vYP_[1], vYP_[2], vKP_[0],
                     vKP_[1], vKP_[2], vMu_[0], vMu_[1], vMu_[2], vMu_[3], gA_, gR_[0], gR_[1], gAcc_,
                     gAP_, gRP_[0], gRP_[1], gAccP_, adj_, vstats_, Mt_, agwc_, agmup_, ad1_, agam_, agamI_,
                     agzp_, agzpp_, agz_, agmu_, agb_, aquirk_, gAD_, gAccD_, gRD_[0], gRD_[1], vYD_[0], vYD_[1],
                     vYD_[2], adjRing_, adjSave_, adjX_, adjSaveZ_, adjSaveP_};
    for (void* p : vbufs)
        if (p) (void)hipFree(p);
    for (auto& kv : traj_) free_record(kv.second);
    for (void* p : {(void*)segZ_, (void*)segP_, (void*)segY_, (void*)segYP_, (void*)segMu_})
        if (p) (void)hipFree(p);
    for (hipEvent_t e : event_pool_) (void)hipEventDestroy(e);
    if (rec_stream_) (void)hipStreamSynchronize(rec_stream_);
    for (hipEvent_t e : rec_packed_) (void)hipEventDestroy(e);
    for (hipEvent_t e : rec_copied_) (void)hipEventDestroy(e);
    if (rec_ring_) (void)hipFree(rec_ring_);
    if (rec_host_) (void)hipHostFree(rec_host_);
    if (rec_stream_) (void)hipStreamDestroy(rec_stream_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

template <typename K, typename... Args>
void Solver::go(K kernel, dim3 grid, dim3 block, unsigned shm, Args... args) {
    if (ev_b_) {
        hipExtLaunchKernelGGL(kernel, grid, block, shm, stream_, ev_a_, ev_b_, 0, args...);
        ev_a_ = nullptr;   // a later launch of the same timed group records only the stop event
    } else {
        hipLaunchKernelGGL(kernel, grid, block, shm, stream_, args...);
    }
}

// f makes one or more kernel launches through go(): the event pair brackets them all, from the first
// kernel's start to the last one's end
template <class F>
void Solver::launch(const char* tag, double flops, double bytes, F&& f, int iter) {
    if (!ktime_) {
        f();
        return;
    }
    while (event_pool_.size() < event_next_ + 2) {
        hipEvent_t e;
        SFH_HIP(hipEventCreate(&e));
        event_pool_.push_back(e);
    }
    hipEvent_t a = event_pool_[event_next_++], b = event_pool_[event_next_++];
    ev_a_ = a;
    ev_b_ = b;
    f();
    ev_a_ = ev_b_ = nullptr;
    timed_.push_back({tag, a, b, flops, bytes, iter, true});
}

void Solver::set_kernel_timing(bool on) {
    SFH_HIP(hipStreamSynchronize(stream_));
    ktime_ = on;
    timed_.clear();
    event_next_ = 0;
}

Solver::KernelReport Solver::kernel_report(const std::string& tag) {
    SFH_HIP(hipStreamSynchronize(stream_));
    KernelReport r;
    for (const Timed& t : timed_) {
        if (tag != t.tag || !t.active) continue;
        float ms = 0.f;
        SFH_HIP(hipEventElapsedTime(&ms, t.a, t.b));
        r.launches += 1;
        r.ms += ms;
        r.flops += t.flops;
        r.bytes += t.bytes;
    }
    return r;
}

GmArgs Solver::gm_args(int stage) const {
    GmArgs a{};
    a.M = M_;
    a.M32 = M32_;
    a.b = b_;
    a.x = gmx0_;
    a.sol = mu_;
    a.V = gmV_;
    a.w = gmw_;
    a.part = gmpart_;
    a.small = gmsmall_;
    a.ctl = gmctl_;
    a.part_stride = gm_part_stride_;
    a.small_stride = gm_small_stride_;
    a.m = seq_len_ > 0 ? seq_len_ : gm_m_;
    a.total = T_;
    a.tol = opt_.tol;
    a.hist = gmhist_;
    a.hvalid = gmhvalid_;
    a.stage = stage;
    return a;
}

// Arnoldi steps of the GMRES solve of all surfaces at once (gridDim.z = surface).  The initial
// residual comes from the assembly kernel and the final step applies the correction x += V y,
// so only the fixed-length sequence of steps is launched here.  With two launches or more the
// last one is gm_close_mu_kernel, which also forms mu' (returns true: no mu_circ_kernel).
// g1 (a sequence of two launches): only the product launch; velocity_gm1_kernel finishes the solve.
// Mixed precision runs the products over the FP32 matrix (4 bytes per entry).
bool Solver::enqueue_gmres(const GmArgs& a, bool g1, bool force64) {
    const bool f32 = opt_.precision == kMixed && !force64;
    double nsq = 0.0;
    for (int v : n_) nsq += (double)v * v;
    const dim3 grid(gm_nwg_, 1, S_);
    const size_t first = timed_.size();
    const bool fused = a.m >= 2;
    const int nprod = fused ? a.m - 1 : a.m;
    for (int j = 0; j < nprod; ++j) {
        // finishing step j - 1 (the basis rebuilt from w_{j-1} and V[0..j-1]) + the product with
        // the stored matrix + the partial dots
        launch(
            a.check_first ? "gm_refine" : "gm_iter", 2.0 * nsq + 4.0 * (j + 1) * T_,
            (f32 ? 4.0 : 8.0) * nsq + 8.0 * (j + 3) * T_,
            [&] {
                if (f32)
                    go(gm_iter_kernel<float>, grid, dim3(256), 0, a,
                                       (const SurfDesc*)d_desc_, j);
                else
                    go(gm_iter_kernel<double>, grid, dim3(256), 0, a,
                                       (const SurfDesc*)d_desc_, j);
            },
            j);
    }
    if (fused && !g1) {
        // finishing the last step + x += V y + mu' = D1 mu: the rebuild and the circulant product
        launch("gm_close_mu", 4.0 * a.m * T_ + 4.0 * nsq, 8.0 * (a.m + 2) * T_ + 24.0 * nsq / kCircRowsPerWG, [&] {
            go(gm_close_mu_kernel, dim3((maxN_ + kCircRowsPerWG - 1) / kCircRowsPerWG, S_), dim3(256),
                               0, a, (const SurfDesc*)d_desc_, (const double*)d_circ_, mup_, a.m - 1);
        });
    }
    if (ktime_ && !a.check_first) {
        // launches 0 .. iters-1 ran a product ("gm_iter"); a gm_iter_kernel launch iters only
        // finished the last step and stored the solution ("gm_close": no product); later ones
        // returned at once.  The fused last launch ("gm_close_mu") always forms mu'.
        std::vector<int> ctl((size_t)kGmCtlInts * S_);
        SFH_HIP(hipStreamSynchronize(stream_));
        SFH_HIP(hipMemcpy(ctl.data(), gmctl_, ctl.size() * sizeof(int), hipMemcpyDeviceToHost));
        int iters = 0;
        for (int s = 0; s < S_; ++s) iters = std::max(iters, ctl[(size_t)kGmCtlInts * s + 2]);
        if (g1) {   // the velocity kernel finishes the step: launch 0 ran a product unless it ended the solve
            iters = 0;
            for (int s = 0; s < S_; ++s)
                if (ctl[(size_t)kGmCtlInts * s] == 0) iters = 1;
        }
        for (size_t k = first; k < timed_.size(); ++k) {
            if (timed_[k].tag != std::string("gm_iter")) continue;
            if (timed_[k].iter == iters) {
                timed_[k].tag = "gm_close";
                timed_[k].flops = 4.0 * (iters + 1) * T_;
                timed_[k].bytes = 8.0 * (iters + 1) * T_;
            } else if (timed_[k].iter > iters) {
                timed_[k].active = false;
            }
        }
    }
    return fused;
}

// M mu = b for every surface; M, b (and its working copy) are ready in stream order.  Returns
// true when mu' has been formed too.
//   Mixed precision (GMRES): the first solve runs over the FP32 matrix from r0 = b - M x0 formed in
//   FP64 by the assembly; then the FP64 residual of its solution is formed matrix-free and a
//   refinement solve over the FP32 matrix starts from it (it ends at launch 0 when the residual is
//   already within the tolerance, the usual case).  Only the refinement solve stores the stage's
//   warm-start history.
bool Solver::enqueue_solve(const double2* Z, int stage, bool g1) {
    if (opt_.solver == kSolveGMRES) {
        if (opt_.precision != kMixed) return enqueue_gmres(gm_args(stage), g1);
        GmArgs first = gm_args(stage);
        first.stage = 0;   // no history: the refinement solve stores the solution
        enqueue_gmres(first);
        GmArgs refine = gm_args(stage);
        double nsq = 0.0;
        for (int v : n_) nsq += (double)v * v;
        launch("residual", (ph_.image ? 2.0 : 1.0) * nsq, 16.0 * T_, [&] {
            SFH_WITH_ROWS(rows_, go(residual64_kernel<R>, dim3((maxN_ + R - 1) / R, S_), dim3(256), 0, d_desc_,
                                    (const NodeTrig*)trig_, Z, Zp_, Zpp_, ph_, refine));
        });
        refine.check_first = 1;
        return enqueue_gmres(refine);
    }
    auto run = [&](int s0, int count, int n) {
        LuArgs a;
        a.A = M_ + desc_[s0].moff;
        a.work = work_ + desc_[s0].voff;
        a.y = y_ + desc_[s0].voff;
        a.x = mu_ + desc_[s0].voff;
        a.uinv = uinv_ + uoff_[s0];
        a.flags = flags_ + foff_[s0];
        a.status = status_ + 2 * s0;
        a.n = n;
        for (int k0 = 0; k0 < n; k0 += kNB) {
            const int kb = std::min(kNB, n - k0);
            const int rem = n - k0 - kb;
            const int nrb = (rem + kNB - 1) / kNB;
            // panel p: L21 and U12 (2 kb^2 rem) and the diagonal block (2 kb^3 class); panel p-1's
            // update of the whole trailing matrix from column k0 on: 2 (n - k0)^2 64
            const double pf = 4.0 * kb * kb * rem + 2.0 * kb * kb * kb;
            const double uf = k0 > 0 ? 2.0 * (n - k0) * (double)(n - k0) * kNB : 0.0;
            const double ub = k0 > 0 ? 8.0 * (2.0 * (n - k0) * (double)(n - k0) + 2.0 * (n - k0) * kNB) : 0.0;
            const int grid = 1 + 2 * nrb + (k0 > 0 ? nrb * nrb : 0);
            launch("lu_step", (double)count * (pf + uf), 8.0 * count * (2.0 * kb * rem + kb * kb) + count * ub, [&] {
                go(lu_step_kernel, dim3(grid, 1, count), dim3(256), 0, a, k0);
            });
        }
        const int P = (n + kNB - 1) / kNB;
        launch("lu_backsolve", (double)count * n * n, 8.0 * count * 0.5 * n * n, [&] {
            go(lu_backsolve_kernel, dim3(P, 1, count), dim3(256), 0, a);
        });
    };
    if (uniform_) {
        run(0, S_, n_[0]);
    } else {
        for (int s = 0; s < S_; ++s) run(s, 1, n_[s]);
    }
    // systems whose interchange-free check failed: re-assemble and solve with GEPP
    go(assemble64_kernel<kPairRows>, dim3((maxN_ + kPairRows - 1) / kPairRows, S_), dim3(256), 0, d_desc_,
       (const NodeTrig*)trig_, Z, Zp_, Zpp_, M_, ph_, (const int*)status_, GmArgs{}, 0);
    go(gepp_fallback_kernel, dim3(S_), dim3(256), 0, d_desc_, M_, b_, mu_, piv_,
                       status_, counters_);
    return false;
}

// One launch for the whole right-hand side (and stage update) of every surface.
void Solver::enqueue_batched(const double2* Z, const double* P, const StageCtl& ctl) {
    BatchArgs a{};
    a.sd = d_desc_;
    a.order = order_;
    a.tw_all = d_tw_;
    a.Zin = Z;
    a.Pin = P;
    a.M = M_;
    a.Zp = Zp_;
    a.Zpp = Zpp_;
    a.mup = mup_;
    a.b = b_;
    a.mu = mu_;
    a.stats = bstats_;
    a.tol = opt_.tol;
    a.mixed = opt_.precision == kMixed ? 1 : 0;
    a.cycles = std::max(1, (opt_.max_iter + 7) / 8);   // kBK = 8 vectors per cycle
    a.max_refine = 6;
    a.nsurf = S_;
    a.total = T_;
    a.hist = bhist_;
    a.hvalid = bhvalid_;
    double nsq = 0.0;
    for (int v : n_) nsq += (double)v * v;
    const double img = ph_.image ? 2.0 : 1.0;
    const double mbytes = (a.mixed ? 4.0 : 8.0) * nsq;
    // algorithmic work per launch: cot pairs of assembly + velocity (and one more pass per
    // refinement residual), matrix written once and read once per GMRES product
    launch("stage", 2.0 * img * nsq, 2.0 * mbytes, [&] {
        go(batched_stage_kernel, dim3(S_), dim3(kBatchThreads), 0, a, ph_, ctl);
    });
    SFH_HIP(hipGetLastError());
}

void Solver::enqueue_rhs(const double2* Z, const double* P, const StageCtl& ctl_in) {
    StageCtl ctl = ctl_in;
    ctl.Pin = P;
    ctl.energy = track_energy_ ? energy_ : nullptr;
    ctl.nonfinite = counters_ + 1;
    if (opt_.solver == kSolveBatched) {
        enqueue_batched(Z, P, ctl);
        return;
    }
    const int gm_on = opt_.solver == kSolveGMRES ? 1 : 0;
    const GmArgs ga = gm_on ? gm_args(ctl.stage) : GmArgs{};
    double nsq = 0.0;
    for (int v : n_) nsq += (double)v * v;
    const double img = ph_.image ? 2.0 : 1.0;
    const dim3 cgrid((maxN_ + kCircRowsPerWG - 1) / kCircRowsPerWG, S_);
    // circulant derivatives: 7 n^2 multiply-adds; bytes: the samples and kernels each workgroup
    // stages (24 B per node and 24 B per kernel entry)
    launch("deriv", 14.0 * nsq, 48.0 * nsq / kCircRowsPerWG, [&] {
        go(zphi_circ_kernel, cgrid, dim3(256), 0, d_desc_, (const double*)d_circ_, Z, P,
                           Zp_, Zpp_, b_, work_, ph_, status_, trig_, ga, gm_on);
    });
    // assembly: flops counted as cot evaluations (the transcendental-bound unit), bytes = M written
    if (gm_on && opt_.precision == kMixed) {
        launch("assemble", img * nsq, 4.0 * nsq, [&] {
            SFH_WITH_ROWS(rows_, go(assemble32_kernel<R>, dim3((maxN_ + R - 1) / R, S_), dim3(256), 0, d_desc_,
                                    (const NodeTrig*)trig_, Z, Zp_, Zpp_, M32_, ph_, ga));
        });
    } else {
        launch("assemble", img * nsq, 8.0 * nsq, [&] {
            SFH_WITH_ROWS(rows_, go(assemble64_kernel<R>, dim3((maxN_ + R - 1) / R, S_), dim3(256), 0, d_desc_,
                                    (const NodeTrig*)trig_, Z, Zp_, Zpp_, M_, ph_, (const int*)nullptr, ga, gm_on));
        });
    }
    // an FP64 GMRES sequence of two launches (one Arnoldi step) ends in the velocity kernel; not
    // with the rho quirk, whose node-1 term needs mu' outside a workgroup's rows
    const bool quirk = ph_.kind == 0 && ph_.rho_quirk && ph_.rho != 0.0;
    const bool g1 = gm_on && opt_.precision != kMixed && ga.m == 2 && !quirk;
    if (!enqueue_solve(Z, ctl.stage, g1)) {
        launch("mu_deriv", 4.0 * nsq, 24.0 * nsq / kCircRowsPerWG, [&] {
            go(mu_circ_kernel, cgrid, dim3(256), 0, d_desc_, (const double*)d_circ_,
                               (const double*)mu_, mup_);
        });
    }
    if (g1) {
        launch("velocity", img * nsq, 96.0 * T_, [&] {
            SFH_WITH_ROWS(rows_, go(velocity_gm1_kernel<R>, dim3((maxN_ + R - 1) / R, S_), dim3(256), 0, d_desc_,
                                    (const NodeTrig*)trig_, Z, Zp_, Zpp_, mup_, ph_, ctl, ga, (const double*)d_circ_));
        });
    } else {
        launch("velocity", img * nsq, 96.0 * T_, [&] {
            SFH_WITH_ROWS(rows_, go(velocity_stage_kernel<R>, dim3((maxN_ + R - 1) / R, S_), dim3(256), 0, d_desc_,
                                    (const NodeTrig*)trig_, Z, Zp_, Zpp_, mu_, mup_, ph_, ctl));
        });
    }
    SFH_HIP(hipGetLastError());
}

void Solver::enqueue_step(double dt, const StageCapture* cap) {
    const double half = dt * 0.5, sixth = dt / 6.0;   // setTimeStep, AutonomousRungeKuttaStepper.cuh:31-36
    StageCtl c{};
    c.Z0 = Z0_;
    c.P0 = P0_;
    for (int i = 0; i < 3; ++i) {
        c.kZ[i] = kZ_[i];
        c.kP[i] = kP_[i];
    }
    c.qup = nullptr;
    // timed drive: stages at t_n, t_n + h/2, t_n + h/2, t_n + h (RK4_Time_Dependent.cuh:173,237,263,270);
    // the first saves the delayed intensity and moves the saved time to t_n
    c.stage = 1; c.alpha = half; c.Zout = Za_; c.Pout = Pa_;
    set_delayed_ctl(c, D0_, Da_);
    timed_coefficients(c, drive_t_, 1);
    enqueue_rhs(Z0_, P0_, c);
    // the stage's mu and the state it produced, for the adjoint's re-run (never inside a captured graph)
    auto keep = [&](int st, const double2* Zs, const double* Ps) {
        if (!cap) return;
        SFH_HIP(hipMemcpyAsync(cap->mu[st], mu_, T_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
        if (st < 3) {
            SFH_HIP(hipMemcpyAsync(cap->Y[st], Zs, T_ * sizeof(double2), hipMemcpyDeviceToDevice, stream_));
            SFH_HIP(hipMemcpyAsync(cap->YP[st], Ps, T_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
        }
    };
    keep(0, Za_, Pa_);
    if (ph_.timed) drive_prev_ = drive_t_;
    c.stage = 2; c.alpha = half; c.Zout = Zb_; c.Pout = Pb_;
    set_delayed_ctl(c, Da_, Db_);
    timed_coefficients(c, drive_t_ + half, 0);
    enqueue_rhs(Za_, Pa_, c);
    keep(1, Zb_, Pb_);
    c.stage = 3; c.alpha = dt; c.Zout = Za_; c.Pout = Pa_;
    set_delayed_ctl(c, Db_, Da_);
    timed_coefficients(c, drive_t_ + half, 0);
    enqueue_rhs(Zb_, Pb_, c);
    keep(2, Za_, Pa_);
    c.stage = 4; c.alpha = sixth; c.Zout = nullptr; c.Pout = nullptr;
    set_delayed_ctl(c, Da_, nullptr);
    timed_coefficients(c, drive_t_ + dt, 0);
    enqueue_rhs(Za_, Pa_, c);
    keep(3, nullptr, nullptr);
    if (ph_.timed) drive_t_ += dt;   // currentTime += timeStep (RK4_Time_Dependent.cuh:408)
}

// calculate_new_delayed_intensity (DelayedIntensityTerm.cuh:20-30) at stage time ts: D = I when ts
// is the saved time, else a Dsaved + beta tau (1 - a) I with a = exp(-(ts - prev) / tau).  Every node
// of the stage takes the same saved time: the reference stores it from every thread of the saving
// stage while other threads may still read it, so some of its nodes see ts == prev; here the time
// moves only after the stage.
void Solver::timed_coefficients(StageCtl& c, double ts, int save) const {
    c.dca = c.dcb = 0.0;
    c.dsave = 0;
    if (!ph_.timed) return;
    if (ts == drive_prev_) {
        c.dcb = 1.0;
    } else {
        const double a = std::exp(-(ts - drive_prev_) / ph_.tau);
        c.dca = a;
        c.dcb = ph_.beta * ph_.tau * (1.0 - a);
    }
    c.dsave = save;
}

void Solver::set_time(double t) {
    if (!std::isfinite(t)) throw ArgError("time must be finite");
    time_ = drive_t_ = drive_prev_ = t;
}

// The delayed-intensity fields of a stage (null without the augmented drive).
void Solver::set_delayed_ctl(StageCtl& c, const double* Din, double* Dout) const {
    if (ph_.timed) {   // the saved delayed intensity, read by every stage, written by the first
        c.Din = D0_;
        c.Dout = D0_;
        c.D0 = nullptr;
        c.kD[0] = c.kD[1] = c.kD[2] = nullptr;
        return;
    }
    if (!ph_.augmented) {
        c.Din = nullptr;
        c.D0 = c.Dout = nullptr;
        c.kD[0] = c.kD[1] = c.kD[2] = nullptr;
        return;
    }
    c.Din = Din;
    c.D0 = D0_;
    c.Dout = Dout;
    for (int i = 0; i < 3; ++i) c.kD[i] = kD_[i];
}

void Solver::capture(double dt, int len) {
    // steps replayed from the old graph may still be queued (sfh_step returns without waiting,
    // and a call with another dt re-captures): drain the stream before destroying it
    SFH_HIP(hipStreamSynchronize(stream_));
    if (exec_) {
        SFH_HIP(hipGraphExecDestroy(exec_));
        exec_ = nullptr;
    }
    if (graph_) {
        SFH_HIP(hipGraphDestroy(graph_));
        graph_ = nullptr;
    }
    seq_len_ = len;
    SFH_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    try {
        enqueue_step(dt);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(stream_, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    SFH_HIP(hipStreamEndCapture(stream_, &graph_));
    SFH_HIP(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
    graph_dt_ = dt;
    graph_len_ = len;
}

void Solver::calibrate() {
    if (++hist_steps_ < kCalibSteps) return;
    // the first step with the full history: its solves (all four stages) bound the Arnoldi steps
    // of the following ones.  L launches run up to L - 1 steps.  Solves that need 1 or 2 steps
    // start from guesses whose error is at the level of the solve tolerance (the extrapolation's
    // truncation error is smaller), and get no spare launch; slower ones, whose step count varies
    // more from solve to solve, get one spare step.  A chunk in which a solve still runs out of
    // launches is run again with the length doubled.
    sync();
    const int need = fill_iters_ + 1;
    trim_ = std::max(2, std::min(gm_m_, fill_iters_ >= 3 ? need + 1 : need));
    if (opt_.force_trim > 0) trim_ = std::max(2, std::min(gm_m_, opt_.force_trim));
    chunk_ = 2;
}

void Solver::save_checkpoint() {
    SFH_HIP(hipMemcpyAsync(ckZ_, Z0_, T_ * sizeof(double2), hipMemcpyDeviceToDevice, stream_));
    SFH_HIP(hipMemcpyAsync(ckP_, P0_, T_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    SFH_HIP(hipMemcpyAsync(ckMu_, mu_, T_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    SFH_HIP(hipMemcpyAsync(ckHist_, gmhist_, 4 * kHistDepth * (size_t)T_ * sizeof(double),
                           hipMemcpyDeviceToDevice, stream_));
    SFH_HIP(hipMemcpyAsync(ckHv_, gmhvalid_, 4 * sizeof(int), hipMemcpyDeviceToDevice, stream_));
    if (ph_.augmented || ph_.timed)
        SFH_HIP(hipMemcpyAsync(ckD_, D0_, T_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    ck_drive_t_ = drive_t_;
    ck_drive_prev_ = drive_prev_;
}

void Solver::restore_checkpoint() {
    SFH_HIP(hipMemcpyAsync(Z0_, ckZ_, T_ * sizeof(double2), hipMemcpyDeviceToDevice, stream_));
    SFH_HIP(hipMemcpyAsync(P0_, ckP_, T_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    SFH_HIP(hipMemcpyAsync(mu_, ckMu_, T_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    SFH_HIP(hipMemcpyAsync(gmhist_, ckHist_, 4 * kHistDepth * (size_t)T_ * sizeof(double),
                           hipMemcpyDeviceToDevice, stream_));
    SFH_HIP(hipMemcpyAsync(gmhvalid_, ckHv_, 4 * sizeof(int), hipMemcpyDeviceToDevice, stream_));
    if (ph_.augmented || ph_.timed)
        SFH_HIP(hipMemcpyAsync(D0_, ckD_, T_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    drive_t_ = ck_drive_t_;
    drive_prev_ = ck_drive_prev_;
}

void Solver::run_steps(double dt, int len, int count) {
    // the timed drive's stage coefficients change from step to step: launched directly
    if (opt_.use_graph && !ktime_ && !ph_.timed) {
        if (!exec_ || graph_dt_ != dt || graph_len_ != len) capture(dt, len);
        for (int b = 0; b < count; ++b) SFH_HIP(hipGraphLaunch(exec_, stream_));
    } else {
        seq_len_ = len;
        for (int b = 0; b < count; ++b) enqueue_step(dt);
    }
}

void Solver::step(double dt, int nsteps) {
    if (nsteps < 0) throw ArgError("negative step count");
    int i = 0;
    while (i < nsteps) {
        if (opt_.solver != kSolveGMRES || (hist_steps_ >= kCalibSteps && trim_ >= gm_m_)) {
            // no trimming (LU, batched, or a GMRES sequence back at the cap): one batch
            run_steps(dt, gm_m_, nsteps - i);
            i = nsteps;
        } else if (hist_steps_ < kCalibSteps) {
            // the warm-start history fills: full-length sequences, one step at a time
            run_steps(dt, gm_m_, 1);
            ++i;
            calibrate();
        } else {
            // a trimmed chunk, kept only when none of its solves ran out of launches
            const int chunk = std::min(chunk_, nsteps - i);
            const int raw0 = unc_raw_;
            save_checkpoint();
            run_steps(dt, trim_, chunk);
            sync();
            if (unc_raw_ > raw0) {
                restore_checkpoint();
                unc_discard_ += unc_raw_ - raw0;
                stats_.unconverged = unc_raw_ - unc_discard_;
                stats_.redone_steps += chunk;
                trim_ = std::min(gm_m_, 2 * trim_);
                chunk_ = 2;
            } else {
                i += chunk;
                chunk_ = std::min(kCheckSteps, 2 * chunk_);
            }
        }
    }
    stats_.steps += nsteps;
    stats_.rhs_evals += 4LL * nsteps;
    for (int k = 0; k < nsteps; ++k) time_ += dt;   // currentTime += dt per step (AutonomousRungeKuttaStepper.cuh:426)
}

__global__ __launch_bounds__(256) void record_frame_kernel(const double2* __restrict__ Z, const double* __restrict__ P,
                                                           const double* __restrict__ D, double* __restrict__ dst,
                                                           int total) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const double2 z = Z[i];
    dst[i] = z.x;
    dst[total + i] = z.y;
    dst[2 * (long long)total + i] = P[i];
    if (D) dst[3 * (long long)total + i] = D[i];
}

void Solver::step_record(double dt, int nsteps, int every, double* states, double* times) {
    if (nsteps < 0 || every < 1) throw ArgError("nsteps must be >= 0 and every >= 1");
    if (nsteps / every > 0 && (!states || !times)) throw ArgError("null output array");
    const size_t frame = (ph_.augmented ? 4 : 3) * (size_t)T_;   // [x, y, phi(, D)]
    if (rec_ring_ && rec_frame_ != frame) {   // the frame size changed (augmented drive): a new ring
        SFH_HIP(hipStreamSynchronize(rec_stream_));
        SFH_HIP(hipFree(rec_ring_));
        SFH_HIP(hipHostFree(rec_host_));
        rec_ring_ = nullptr;
        rec_host_ = nullptr;
    }
    if (!rec_ring_) {
        rec_frame_ = frame;
        // ring depth: up to 32 frames, at most 64 MiB of frames
        rec_slots_ = (int)std::max<size_t>(2, std::min<size_t>(32, (64u << 20) / (frame * sizeof(double))));
        rec_ring_ = dalloc<double>(frame * rec_slots_);
        SFH_HIP(hipHostMalloc((void**)&rec_host_, frame * rec_slots_ * sizeof(double), hipHostMallocDefault));
        if (!rec_stream_) SFH_HIP(hipStreamCreateWithFlags(&rec_stream_, hipStreamNonBlocking));
        while ((int)rec_packed_.size() < rec_slots_) {
            hipEvent_t a, b;
            SFH_HIP(hipEventCreateWithFlags(&a, hipEventDisableTiming));
            SFH_HIP(hipEventCreateWithFlags(&b, hipEventDisableTiming));
            rec_packed_.push_back(a);
            rec_copied_.push_back(b);
        }
    }
    const int frames = nsteps / every;
    auto drain = [&](int f) {   // frame f's copy has been issued: wait for it, hand it over
        const int slot = f % rec_slots_;
        SFH_HIP(hipEventSynchronize(rec_copied_[slot]));
        std::memcpy(states + (size_t)f * frame, rec_host_ + (size_t)slot * frame, frame * sizeof(double));
    };
    for (int f = 0; f < frames; ++f) {
        step(dt, every);   // a GMRES chunk that was redone never reaches a frame: step() keeps only checked steps
        times[f] = time_;
        const int slot = f % rec_slots_;
        if (f >= rec_slots_) drain(f - rec_slots_);   // the slot's previous frame is on the host
        double* dst = rec_ring_ + (size_t)slot * frame;
        hipLaunchKernelGGL(record_frame_kernel, dim3((T_ + 255) / 256), dim3(256), 0, stream_, Z0_, P0_,
                           (const double*)(ph_.augmented ? D0_ : nullptr), dst, T_);
        SFH_HIP(hipGetLastError());
        SFH_HIP(hipEventRecord(rec_packed_[slot], stream_));
        SFH_HIP(hipStreamWaitEvent(rec_stream_, rec_packed_[slot], 0));
        SFH_HIP(hipMemcpyAsync(rec_host_ + (size_t)slot * frame, dst, frame * sizeof(double), hipMemcpyDeviceToHost,
                               rec_stream_));
        SFH_HIP(hipEventRecord(rec_copied_[slot], rec_stream_));
    }
    for (int f = std::max(0, frames - rec_slots_); f < frames; ++f) drain(f);
    if (nsteps > frames * every) step(dt, nsteps - frames * every);
    sync();
}

__global__ __launch_bounds__(256) void state_split_kernel(const double2* __restrict__ state, double2* __restrict__ Z,
                                                          double* __restrict__ P, int total) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    Z[i] = state[i];
    P[i] = state[total + i].x;
}

__global__ __launch_bounds__(256) void state_join_kernel(const double2* __restrict__ Z, const double* __restrict__ P,
                                                         double2* __restrict__ state, int total) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    state[i] = Z[i];
    state[total + i] = make_double2(P[i], 0.0);
}

// A new trajectory: the non-finite flag and every solver vector a solve may read before writing it
// (the last mu, the initial guess, the GMRES basis, products, partial sums and small state) start
// from zero, so that the same state gives the same results whatever the handle ran before, and a
// trajectory that blew up leaves nothing non-finite behind.  Not part of any captured graph.
void Solver::clear_solver_state() {
    SFH_HIP(hipMemsetAsync(counters_ + 1, 0, sizeof(int), stream_));
    SFH_HIP(hipMemsetAsync(mu_, 0, T_ * sizeof(double), stream_));
    if (opt_.solver == kSolveGMRES) {
        SFH_HIP(hipMemsetAsync(gmx0_, 0, T_ * sizeof(double), stream_));
        SFH_HIP(hipMemsetAsync(gmV_, 0, (size_t)(gm_m_ + 1) * T_ * sizeof(double), stream_));
        SFH_HIP(hipMemsetAsync(gmw_, 0, 2 * (size_t)T_ * sizeof(double), stream_));
        SFH_HIP(hipMemsetAsync(gmpart_, 0, gm_part_stride_ * S_ * sizeof(double), stream_));
        SFH_HIP(hipMemsetAsync(gmsmall_, 0, gm_small_stride_ * S_ * sizeof(double), stream_));
    }
}

void Solver::set_state_device(const double2* state) {
    if (!state) throw ArgError("null state");
    hipLaunchKernelGGL(state_split_kernel, dim3((T_ + 255) / 256), dim3(256), 0, stream_, state, Z0_, P0_, T_);
    SFH_HIP(hipGetLastError());
    clear_solver_state();
    if (gmhvalid_) SFH_HIP(hipMemsetAsync(gmhvalid_, 0, 4 * sizeof(int), stream_));   // a new trajectory
    if (bhvalid_) SFH_HIP(hipMemsetAsync(bhvalid_, 0, 4 * S_ * sizeof(int), stream_));
    hist_steps_ = 0;
    time_ = drive_t_ = drive_prev_ = 0.0;
    SFH_HIP(hipStreamSynchronize(stream_));
}

void Solver::get_state_device(double2* state) {
    if (!state) throw ArgError("null state");
    hipLaunchKernelGGL(state_join_kernel, dim3((T_ + 255) / 256), dim3(256), 0, stream_, Z0_, P0_, state, T_);
    SFH_HIP(hipGetLastError());
    sync();
}

void Solver::rhs_device(const double2* state, double2* rhs) {
    if (!state || !rhs) throw ArgError("null state or rhs");
    if (ph_.augmented) throw ArgError("rhs_device takes the two-field state; the augmented drive has three");
    // the stage buffers Za/Pa are free between steps
    hipLaunchKernelGGL(state_split_kernel, dim3((T_ + 255) / 256), dim3(256), 0, stream_, state, Za_, Pa_, T_);
    SFH_HIP(hipGetLastError());
    seq_len_ = gm_m_;   // a standalone solve: full launch sequence
    StageCtl c{};
    c.stage = 0;
    c.kZ[0] = kZ_[0];
    c.kP[0] = kP_[0];
    set_delayed_ctl(c, D0_, nullptr);
    timed_coefficients(c, drive_t_, 0);
    enqueue_rhs(Za_, Pa_, c);
    hipLaunchKernelGGL(state_join_kernel, dim3((T_ + 255) / 256), dim3(256), 0, stream_, kZ_[0], kP_[0], rhs, T_);
    SFH_HIP(hipGetLastError());
    stats_.rhs_evals += 1;
    sync();
}

// Two nodes of one surface at the same point (x mod 2 pi, y) make cot((z_k - z_j)/2) infinite and
// the dipole matrix singular.  Rounded in FMA the pair terms come out huge but finite rather than
// NaN, so the case is caught here, by sorting each surface's nodes.
static void check_distinct_nodes(const double* x, const double* y, const std::vector<int>& sizes) {
    std::vector<std::pair<double, double>> pts;
    std::vector<int> idx;
    int off = 0;
    for (size_t s = 0; s < sizes.size(); ++s) {
        const int n = sizes[s];
        pts.resize(n);
        idx.resize(n);
        for (int i = 0; i < n; ++i) {
            double xr = std::fmod(x[off + i], 2.0 * kPi);
            if (xr < 0.0) xr += 2.0 * kPi;
            if (xr >= 2.0 * kPi) xr = 0.0;
            pts[i] = {xr, y[off + i]};
            idx[i] = i;
        }
        std::sort(idx.begin(), idx.end(), [&](int a, int b) { return pts[a] < pts[b]; });
        for (int i = 1; i < n; ++i)
            if (pts[idx[i]] == pts[idx[i - 1]])
                throw ArgError("coincident nodes " + std::to_string(std::min(idx[i], idx[i - 1])) + " and " +
                               std::to_string(std::max(idx[i], idx[i - 1])) + " of surface " + std::to_string(s) +
                               ": the dipole matrix is singular");
        off += n;
    }
}

void Solver::set_state(const double* x, const double* y, const double* phi) {
    for (int i = 0; i < T_; ++i)
        if (!std::isfinite(x[i]) || !std::isfinite(y[i]) || !std::isfinite(phi[i]))
            throw ArgError("non-finite state at node " + std::to_string(i));
    check_distinct_nodes(x, y, n_);
    std::vector<double2> z(T_);
    for (int i = 0; i < T_; ++i) z[i] = make_double2(x[i], y[i]);
    clear_solver_state();
    SFH_HIP(hipMemcpyAsync(Z0_, z.data(), T_ * sizeof(double2), hipMemcpyHostToDevice, stream_));
    SFH_HIP(hipMemcpyAsync(P0_, phi, T_ * sizeof(double), hipMemcpyHostToDevice, stream_));
    if (gmhvalid_) SFH_HIP(hipMemsetAsync(gmhvalid_, 0, 4 * sizeof(int), stream_));   // a new trajectory
    hist_steps_ = 0;
    time_ = drive_t_ = drive_prev_ = 0.0;
    if (D0_) SFH_HIP(hipMemsetAsync(D0_, 0, T_ * sizeof(double), stream_));
    if (bhvalid_) SFH_HIP(hipMemsetAsync(bhvalid_, 0, 4 * S_ * sizeof(int), stream_));
    SFH_HIP(hipStreamSynchronize(stream_));
}

void Solver::get_state(double* x, double* y, double* phi) {
    std::vector<double2> z(T_);
    SFH_HIP(hipMemcpyAsync(z.data(), Z0_, T_ * sizeof(double2), hipMemcpyDeviceToHost, stream_));
    SFH_HIP(hipMemcpyAsync(phi, P0_, T_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
    SFH_HIP(hipStreamSynchronize(stream_));
    for (int i = 0; i < T_; ++i) {
        x[i] = z[i].x;
        y[i] = z[i].y;
    }
    check_status();
}

void Solver::rhs(double* vx, double* vy, double* dphi, double* mu, double2* zp, double2* zpp, double2* qup) {
    seq_len_ = gm_m_;   // a standalone solve: no extrapolated guess, the full launch sequence
    StageCtl c{};
    c.stage = 0;
    c.kZ[0] = kZ_[0];
    c.kP[0] = kP_[0];
    c.qup = qup ? qup_ : nullptr;
    set_delayed_ctl(c, D0_, nullptr);
    timed_coefficients(c, drive_t_, 0);   // timed drive: at the handle's time, nothing saved
    enqueue_rhs(Z0_, P0_, c);
    std::vector<double2> q(T_);
    SFH_HIP(hipMemcpyAsync(q.data(), kZ_[0], T_ * sizeof(double2), hipMemcpyDeviceToHost, stream_));
    SFH_HIP(hipMemcpyAsync(dphi, kP_[0], T_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
    if (mu) SFH_HIP(hipMemcpyAsync(mu, mu_, T_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
    if (zp) SFH_HIP(hipMemcpyAsync(zp, Zp_, T_ * sizeof(double2), hipMemcpyDeviceToHost, stream_));
    if (zpp) SFH_HIP(hipMemcpyAsync(zpp, Zpp_, T_ * sizeof(double2), hipMemcpyDeviceToHost, stream_));
    if (qup) SFH_HIP(hipMemcpyAsync(qup, qup_, T_ * sizeof(double2), hipMemcpyDeviceToHost, stream_));
    SFH_HIP(hipStreamSynchronize(stream_));
    for (int i = 0; i < T_; ++i) {
        vx[i] = q[i].x;
        vy[i] = q[i].y;
    }
    stats_.rhs_evals += 1;
    check_status();
}

void Solver::set_drive(const Phys& d) {
    if (d.augmented && !d.drive) throw ArgError("the augmented system needs the drive on");
    if (d.timed && (!d.drive || d.augmented)) throw ArgError("the timed drive needs the drive on and no augmented system");
    if (d.drive && !(d.tau != 0.0)) throw ArgError("tau must be nonzero");
    SFH_HIP(hipStreamSynchronize(stream_));
    ph_.drive = d.drive;
    ph_.augmented = d.augmented;
    ph_.timed = d.timed;
    drive_prev_ = drive_t_ = time_;
    ph_.damping = d.damping;
    ph_.strength = d.strength;
    ph_.detuning = d.detuning;
    ph_.gamma = d.gamma;
    ph_.G = d.G;
    ph_.max_intensity = d.max_intensity;
    ph_.x0 = d.x0;
    ph_.sigma = d.sigma;
    ph_.beta = d.beta;
    ph_.tau = d.tau;
    if ((ph_.augmented || ph_.timed) && !D0_) {
        D0_ = dalloc<double>(T_);
        Da_ = dalloc<double>(T_);
        Db_ = dalloc<double>(T_);
        ckD_ = dalloc<double>(T_);
        for (int i = 0; i < 3; ++i) kD_[i] = dalloc<double>(T_);
    }
    if (D0_) SFH_HIP(hipMemset(D0_, 0, T_ * sizeof(double)));
    graph_len_ = -1;   // the step graph holds the physics: re-capture
}

void Solver::set_delayed(const double* D) {
    if (!ph_.augmented && !ph_.timed) throw ArgError("the handle has no augmented or timed drive (sfh_set_drive)");
    SFH_HIP(hipMemcpyAsync(D0_, D, T_ * sizeof(double), hipMemcpyHostToDevice, stream_));
    SFH_HIP(hipStreamSynchronize(stream_));
}

void Solver::get_delayed(double* D) {
    if (!ph_.augmented && !ph_.timed) throw ArgError("the handle has no augmented or timed drive (sfh_set_drive)");
    SFH_HIP(hipMemcpyAsync(D, D0_, T_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
    SFH_HIP(hipStreamSynchronize(stream_));
}

void Solver::get_delayed_rhs(double* dD) {
    if (!ph_.augmented) throw ArgError("the handle has no augmented drive (sfh_set_drive)");
    SFH_HIP(hipMemcpyAsync(dD, kD_[0], T_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
    SFH_HIP(hipStreamSynchronize(stream_));
}

void Solver::track_energies(bool on) {
    SFH_HIP(hipStreamSynchronize(stream_));
    if (on && !energy_) {
        energy_ = dalloc<double>(4 * (size_t)T_);
        SFH_HIP(hipMemset(energy_, 0, 4 * (size_t)T_ * sizeof(double)));
    }
    if (on != track_energy_) graph_len_ = -1;   // the step graph holds the energy pointer: re-capture
    track_energy_ = on;
}

void Solver::energies(double* out) {
    if (!track_energy_) throw ArgError("energies are not tracked (sfh_track_energies)");
    std::vector<double> e(4 * (size_t)T_);
    SFH_HIP(hipMemcpyAsync(e.data(), energy_, e.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
    SFH_HIP(hipStreamSynchronize(stream_));
    // scales of KineticEnergy, GravitationalEnergy / VanDerWaalsEnergy, SurfaceEnergy and VolumeFlux
    // (Energies.cuh:61-128); the potential is multiplied by g (1 in the reference)
    for (int s = 0; s < S_; ++s) {
        double sum[4] = {0.0, 0.0, 0.0, 0.0};
        for (int i = 0; i < n_[s]; ++i)
            for (int c = 0; c < 4; ++c) sum[c] += e[4 * ((size_t)desc_[s].voff + i) + c];
        out[4 * s + 0] = sum[0] * 0.25 / kPi;
        out[4 * s + 1] = ph_.g * (ph_.kind == 0 ? sum[1] * 0.25 * (1.0 + ph_.rho) / kPi
                                                : sum[1] * ph_.depth * ph_.depth / 6.0);
        out[4 * s + 2] = (sum[2] - 2.0 * kPi) * ph_.kappa / (2.0 * kPi);
        out[4 * s + 3] = sum[3] * 0.5 / kPi;
    }
}

void Solver::get_mu(double* mu) {
    SFH_HIP(hipMemcpyAsync(mu, mu_, T_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
    SFH_HIP(hipStreamSynchronize(stream_));
}

void Solver::sync() {
    SFH_HIP(hipStreamSynchronize(stream_));
    check_status();
}

void Solver::check_status() {
    std::vector<int> st(2 * S_);
    int cnt[4];
    SFH_HIP(hipMemcpy(st.data(), status_, 2 * S_ * sizeof(int), hipMemcpyDeviceToHost));
    SFH_HIP(hipMemcpy(cnt, counters_, sizeof(cnt), hipMemcpyDeviceToHost));
    stats_.fallback_solves = cnt[0];
    int timeouts = 0;
    for (int s = 0; s < S_; ++s) timeouts += st[2 * s + 1] != 0;
    stats_.spin_timeouts = timeouts;
    if (opt_.solver == kSolveGMRES) {
        std::vector<int> ctl((size_t)kGmCtlInts * S_);
        SFH_HIP(hipMemcpy(ctl.data(), gmctl_, ctl.size() * sizeof(int), hipMemcpyDeviceToHost));
        int mx = 0, unc = 0, fill = 0;
        for (int s = 0; s < S_; ++s) {
            GmCtl c;
            std::memcpy(&c, ctl.data() + (size_t)kGmCtlInts * s, sizeof(GmCtl));
            mx = std::max(mx, c.iters);
            unc += c.unconverged;
            fill = std::max(fill, c.fill_iters);
        }
        fill_iters_ = fill;
        unc_raw_ = unc;
        unc -= unc_discard_;
        stats_.last_iterations = ctl[2];
        // mixed precision: 1 when surface 0's last refinement solve ran Arnoldi steps
        if (opt_.precision == kMixed) stats_.refinements = ctl[1] > 0 ? 1 : 0;
        stats_.max_iterations = mx;
        stats_.unconverged = unc;
    }
    if (opt_.solver == kSolveBatched) {
        std::vector<int> bs(4 * (size_t)S_);
        SFH_HIP(hipMemcpy(bs.data(), bstats_, bs.size() * sizeof(int), hipMemcpyDeviceToHost));
        int mx = 0, unc = 0;
        for (int s = 0; s < S_; ++s) {
            mx = std::max(mx, bs[4 * s]);
            unc += bs[4 * s + 1];
        }
        stats_.last_iterations = bs[0];
        stats_.max_iterations = mx;
        stats_.unconverged = unc;
        stats_.refinements = bs[2];
    }
    if (timeouts) throw HipError("back-substitution wait timed out; results are invalid");
    if (cnt[1])
        throw NonFiniteError("non-finite right-hand side (NaN or infinite state, or coincident nodes); set a new state");
}

Stats Solver::stats() {
    sync();
    stats_.sequence_length = (opt_.solver == kSolveGMRES && hist_steps_ >= kCalibSteps) ? trim_ : 0;
    return stats_;
}

void Solver::derivative(const double2* in_host, double2* out_host, bool second, double scale) {
    SFH_HIP(hipMemcpyAsync(Za_, in_host, T_ * sizeof(double2), hipMemcpyHostToDevice, stream_));
    const int lds_tw = (2 * (size_t)maxN_ * sizeof(double2) <= 144 * 1024) ? 1 : 0;
    const size_t lds = (size_t)maxN_ * sizeof(double2) * (lds_tw ? 2 : 1);
    go(plain_deriv_kernel, dim3(S_), dim3(kDerivThreads), lds, d_desc_, d_tw_, Za_, Zb_,
                       second ? 1 : 0, scale, lds_tw);
    SFH_HIP(hipGetLastError());
    SFH_HIP(hipMemcpyAsync(out_host, Zb_, T_ * sizeof(double2), hipMemcpyDeviceToHost, stream_));
    SFH_HIP(hipStreamSynchronize(stream_));
}

}  // namespace sfh

namespace sfh {

// Solves A x = b (n x n, row-major, host arrays) with the device LU path used by the stepper:
// interchange-free blocked LU, and the GEPP kernel when a multiplier exceeds 1.  Returns 1 in
// *interchanged when the GEPP kernel ran.  Used to test the solver on arbitrary matrices.
void debug_lu_solve(int n, const double* A, const double* b, double* x, int* interchanged) {
    if (n < 1) throw ArgError("n must be positive");
    hipStream_t st;
    SFH_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    const int P = (n + kNB - 1) / kNB;
    double *dA, *dA0, *dwork, *dy, *dx, *duinv, *db;
    int *dflags, *dstatus, *dpiv, *dcnt;
    SurfDesc desc{};
    desc.n = n;
    SurfDesc* ddesc;
    SFH_HIP(hipMalloc(&dA, sizeof(double) * n * n));
    SFH_HIP(hipMalloc(&dA0, sizeof(double) * n * n));
    SFH_HIP(hipMalloc(&dwork, sizeof(double) * n));
    SFH_HIP(hipMalloc(&db, sizeof(double) * n));
    SFH_HIP(hipMalloc(&dy, sizeof(double) * n));
    SFH_HIP(hipMalloc(&dx, sizeof(double) * n));
    SFH_HIP(hipMalloc(&duinv, sizeof(double) * P * kNB * kNB));
    SFH_HIP(hipMalloc(&dflags, sizeof(int) * P));
    SFH_HIP(hipMalloc(&dstatus, sizeof(int) * 2));
    SFH_HIP(hipMalloc(&dpiv, sizeof(int) * n));
    SFH_HIP(hipMalloc(&dcnt, sizeof(int) * 4));
    SFH_HIP(hipMalloc(&ddesc, sizeof(SurfDesc)));
    SFH_HIP(hipMemcpy(ddesc, &desc, sizeof(desc), hipMemcpyHostToDevice));
    SFH_HIP(hipMemcpy(dA, A, sizeof(double) * n * n, hipMemcpyHostToDevice));
    SFH_HIP(hipMemcpy(dA0, A, sizeof(double) * n * n, hipMemcpyHostToDevice));
    SFH_HIP(hipMemcpy(dwork, b, sizeof(double) * n, hipMemcpyHostToDevice));
    SFH_HIP(hipMemcpy(db, b, sizeof(double) * n, hipMemcpyHostToDevice));
    SFH_HIP(hipMemset(dstatus, 0, sizeof(int) * 2));
    SFH_HIP(hipMemset(dcnt, 0, sizeof(int) * 4));
    LuArgs a{dA, dwork, dy, dx, duinv, dflags, dstatus, n};
    for (int k0 = 0; k0 < n; k0 += kNB) {
        const int kb = std::min(kNB, n - k0);
        const int rem = n - k0 - kb;
        const int nrb = (rem + kNB - 1) / kNB;
        hipLaunchKernelGGL(lu_step_kernel, dim3(1 + 2 * nrb + (k0 > 0 ? nrb * nrb : 0), 1, 1), dim3(256), 0, st,
                           a, k0);
    }
    hipLaunchKernelGGL(lu_backsolve_kernel, dim3(P, 1, 1), dim3(256), 0, st, a);
    SFH_HIP(hipGetLastError());
    int status[2];
    SFH_HIP(hipMemcpyAsync(status, dstatus, sizeof(status), hipMemcpyDeviceToHost, st));
    SFH_HIP(hipStreamSynchronize(st));
    if (status[0]) {
        SFH_HIP(hipMemcpyAsync(dA, dA0, sizeof(double) * n * n, hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(gepp_fallback_kernel, dim3(1), dim3(256), 0, st, ddesc, dA, db, dx, dpiv, dstatus, dcnt);
        SFH_HIP(hipGetLastError());
    }
    SFH_HIP(hipMemcpyAsync(x, dx, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    SFH_HIP(hipStreamSynchronize(st));
    if (interchanged) *interchanged = status[0] ? 1 : 0;
    void* bufs[] = {dA, dA0, dwork, db, dy, dx, duinv, dflags, dstatus, dpiv, dcnt, ddesc};
    for (void* p : bufs) (void)hipFree(p);
    (void)hipStreamDestroy(st);
    if (status[1]) throw HipError("back-substitution wait timed out");
}

// ------------------------------------------------------------------ reverse mode

void Solver::ensure_vjp() {
    if (ph_.drive && ph_.timed)
        throw ArgError("the adjoint is implemented for water, the helium film and the autonomous and augmented "
                       "drives (not the time-dependent delayed intensity)");
    if (opt_.solver == kSolveBatched && maxN_ > kBatchMaxN)
        throw ArgError("the batched adjoint takes surfaces of at most 1024 nodes");
    if (vjp_ready_) return;
    if (opt_.solver != kSolveBatched) {   // the multi-workgroup reverse pass
        alloc_gmres(kGmMaxIter);
        Mt_ = dalloc<double>(MT_);
        for (double2** b : {&agwc_, &agmup_, &ad1_, &agam_, &agamI_, &agzp_, &agzpp_, &agz_}) *b = dalloc<double2>(T_);
        agmu_ = dalloc<double>(T_);
        agb_ = dalloc<double>(T_);
        aquirk_ = dalloc<double>((size_t)3 * T_);
        SFH_HIP(hipMemset(agamI_, 0, T_ * sizeof(double2)));
        adjRing_ = dalloc<double>((size_t)4 * kHistDepth * T_);
        adjSave_ = dalloc<double>((size_t)4 * kHistDepth * T_);
        adjX_ = dalloc<double>(T_);
        adjSaveZ_ = dalloc<double2>(T_);
        adjSaveP_ = dalloc<double>(T_);
    }
    for (int i = 0; i < 3; ++i) {
        vY_[i] = dalloc<double2>(T_);
        vK_[i] = dalloc<double2>(T_);
        vYP_[i] = dalloc<double>(T_);
        vKP_[i] = dalloc<double>(T_);
    }
    for (int i = 0; i < 4; ++i) vMu_[i] = dalloc<double>(T_);
    gA_ = dalloc<double2>(T_);
    gAcc_ = dalloc<double2>(T_);
    gAP_ = dalloc<double>(T_);
    gAccP_ = dalloc<double>(T_);
    for (int i = 0; i < 2; ++i) {
        gR_[i] = dalloc<double2>(T_);
        gRP_[i] = dalloc<double>(T_);
    }
    adj_ = dalloc<AdjNode>(T_);
    vstats_ = dalloc<int>(4 * (size_t)S_);
    SFH_HIP(hipMemset(vstats_, 0, 4 * S_ * sizeof(int)));
    vjp_ready_ = true;
}

void Solver::enqueue_vjp(const double2* Z, const double* P, const double* mu, const double2* gAZ, const double* gAP,
                         double ca, const double2* gBZ, const double* gBP, double cb, double2* outZ, double* outP,
                         bool accumulate, int stage, const double* gAD, const double* gBD, double* outD) {
    auto aug = [&] {   // the delayed intensity's terms (augmented drive)
        if (!ph_.augmented || !outD) return;
        go(aug_vjp_kernel, dim3((T_ + 255) / 256), dim3(256), 0, Z, gAP, gBP, gAD, gBD, ca, cb, outZ,
           accumulate ? gAcc_ : nullptr, outD, accumulate ? gAccD_ : nullptr, ph_, T_);
    };
    if (opt_.solver != kSolveBatched) {
        enqueue_vjp_mwg(Z, P, mu, gAZ, gAP, ca, gBZ, gBP, cb, outZ, outP, accumulate, stage);
        aug();
        return;
    }
    VjpArgs a{};
    a.sd = d_desc_;
    a.tw_all = d_tw_;
    a.Zin = Z;
    a.Pin = P;
    a.mu_in = mu;
    a.gA_Z = gAZ;
    a.gA_P = gAP;
    a.gB_Z = gBZ;
    a.gB_P = gBP;
    a.ca = ca;
    a.cb = cb;
    a.out_Z = outZ;
    a.out_P = outP;
    a.acc_Z = accumulate ? gAcc_ : nullptr;
    a.acc_P = accumulate ? gAccP_ : nullptr;
    a.scratch = adj_;
    a.stats = vstats_;
    a.tol = std::min(opt_.tol, 1e-13);
    a.cycles = 6;
    double nsq = 0.0;
    for (int v : n_) nsq += (double)v * v;
    launch("vjp", 3.0 * nsq, 0.0, [&] {
        go(rhs_vjp_kernel, dim3(S_), dim3(kBatchThreads), 0, a, ph_);
    });
    aug();
    SFH_HIP(hipGetLastError());
}

// The reverse pass over many workgroups per surface (adjoint_mwg.hip), at the state (Z, P) with its
// forward mu: the derivatives and mu' of the state, pass A, D1 of the cotangent of mu', pass B, the
// transposed solve M^T gb = gmu (M re-assembled and transposed, the GMRES launch sequence from a zero
// guess, FP64), the matrix terms and pair weights, pass C, and the spectral adjoints.
void Solver::enqueue_vjp_mwg(const double2* Z, const double* P, const double* mu, const double2* gAZ,
                             const double* gAP, double ca, const double2* gBZ, const double* gBP, double cb,
                             double2* outZ, double* outP, bool accumulate, int stage) {
    double nsq = 0.0;
    for (int v : n_) nsq += (double)v * v;
    const double img = ph_.image ? 2.0 : 1.0;
    const dim3 cgrid((maxN_ + kCircRowsPerWG - 1) / kCircRowsPerWG, S_);
    const dim3 pgrid((maxN_ + kAdjRows - 1) / kAdjRows, S_);
    const dim3 wgrid((maxN_ + 3) / 4, S_);
    const dim3 ngrid((maxN_ + 255) / 256, S_);
    launch("vjp_fwd", 14.0 * nsq, 0.0, [&] {
        go(zphi_circ_kernel, cgrid, dim3(256), 0, d_desc_, (const double*)d_circ_, Z, P, Zp_, Zpp_, b_, work_, ph_,
           (int*)nullptr, trig_, GmArgs{}, 0);
        go(mu_circ_kernel, cgrid, dim3(256), 0, d_desc_, (const double*)d_circ_, mu, mup_);
    });
    AdjMArgs a{};
    a.sd = d_desc_;
    a.trig = trig_;
    a.Z = Z;
    a.Zp = Zp_;
    a.Zpp = Zpp_;
    a.mup = mup_;
    a.mu = mu;
    a.gA_Z = gAZ;
    a.gA_P = gAP;
    a.gB_Z = gBZ;
    a.gB_P = gBP;
    a.ca = ca;
    a.cb = cb;
    a.node = adj_;
    a.gwc = agwc_;
    a.gmup = agmup_;
    a.d1 = ad1_;
    a.gmu = agmu_;
    a.gb = agb_;
    a.gam = agam_;
    a.gamI = agamI_;
    a.gzp = agzp_;
    a.gzpp = agzpp_;
    a.gz = agz_;
    a.out_Z = outZ;
    a.out_P = outP;
    a.acc_Z = accumulate ? gAcc_ : nullptr;
    a.acc_P = accumulate ? gAccP_ : nullptr;
    a.quirk = aquirk_;
    a.total = T_;
    const bool quirk = ph_.kind == 0 && ph_.rho != 0.0 && ph_.rho_quirk;
    launch("vjp_pairs", img * 3.0 * nsq, 0.0, [&] {
        go(adj_pass_a_kernel<kAdjRows>, pgrid, dim3(256), 0, a, ph_);
        if (quirk) go(adj_rho_quirk_kernel, dim3((maxN_ + 255) / 256, S_), dim3(256), 0, a);
        go(adj_circ_d1_kernel, wgrid, dim3(256), 0, d_desc_, (const double*)d_circ_, (const double2*)agmup_, ad1_);
        go(adj_pass_b_kernel<kAdjRows>, pgrid, dim3(256), 0, a, ph_);
    });
    // M^T gb = gmu
    launch("vjp_solve", 2.0 * gm_m_ * nsq, 8.0 * gm_m_ * nsq, [&] {
        go(assemble64_kernel<kPairRows>, dim3((maxN_ + kPairRows - 1) / kPairRows, S_), dim3(256), 0, d_desc_,
           (const NodeTrig*)trig_, Z, Zp_, Zpp_, M_, ph_, (const int*)nullptr, GmArgs{}, 0);
        go(transpose_kernel, dim3((maxN_ + 31) / 32, (maxN_ + 31) / 32, S_), dim3(256), 0, d_desc_, (const double*)M_,
           Mt_);
        // inside a trajectory sweep: the guess extrapolated from this stage's solutions of the previous
        // reverse steps, and once calibrated a trimmed sequence; else a zero guess, the full sequence
        const bool warm = adj_warm_ && stage >= 0 && adj_hv_[stage] > 0;
        const int m = (warm && adj_len_ > 0 && adj_hv_[stage] >= kHistDepth) ? adj_len_ : gm_m_;
        if (warm) {
            go(adj_extrapolate_kernel, dim3((T_ + 255) / 256), dim3(256), 0,
               (const double*)(adjRing_ + (size_t)stage * kHistDepth * T_), adj_head_[stage], adj_hv_[stage], adjX_,
               T_);
            go(adj_gm_start_warm_kernel, dim3((maxN_ + 3) / 4, S_), dim3(256), 0, d_desc_, (const double*)Mt_,
               (const double*)agmu_, (const double*)adjX_, gmV_, gmx0_, gmctl_, m);
        } else {
            go(adj_gm_start_kernel, ngrid, dim3(256), 0, d_desc_, (const double*)agmu_, gmV_, gmx0_, gmctl_, m);
        }
        GmArgs g = gm_args(0);
        g.M = Mt_;
        g.M32 = nullptr;
        g.b = agmu_;
        g.x = gmx0_;
        g.sol = agb_;
        g.m = m;
        g.tol = std::min(opt_.tol, 1e-13);
        g.hist = nullptr;
        g.hvalid = nullptr;
        g.stage = 0;
        g.check_first = 0;
        enqueue_gmres(g, false, true);   // leaves s D1 gb in mup_
    });
    if (adj_warm_ && stage >= 0) {   // the solution joins the stage's ring
        const int slot = (adj_head_[stage] + 1) % kHistDepth;
        SFH_HIP(hipMemcpyAsync(adjRing_ + ((size_t)stage * kHistDepth + slot) * T_, agb_, T_ * sizeof(double),
                               hipMemcpyDeviceToDevice, stream_));
        adj_head_[stage] = slot;
        adj_hv_[stage] = std::min(adj_hv_[stage] + 1, kHistDepth);
    }
    if (adj_calib_) {   // the Arnoldi steps this solve took
        std::vector<int> ctl((size_t)kGmCtlInts * S_);
        SFH_HIP(hipStreamSynchronize(stream_));
        SFH_HIP(hipMemcpy(ctl.data(), gmctl_, ctl.size() * sizeof(int), hipMemcpyDeviceToHost));
        for (int s = 0; s < S_; ++s) adj_calib_max_ = std::max(adj_calib_max_, ctl[(size_t)kGmCtlInts * s + 2]);
    }
    launch("vjp_pairs", img * nsq, 0.0, [&] {
        go(adj_gamma_kernel, ngrid, dim3(256), 0, a, ph_);
        go(adj_pass_c_kernel<kAdjRows>, pgrid, dim3(256), 0, a, ph_);
        go(adj_final_kernel, wgrid, dim3(256), 0, a, (const double*)d_circ_, (const double2*)mup_);
    });
    SFH_HIP(hipGetLastError());
}

void Solver::upload_cotangent(const double* gx, const double* gy, const double* gp, double2* Z, double* P) {
    std::vector<double2> z(T_);
    for (int i = 0; i < T_; ++i) z[i] = make_double2(gx[i], gy[i]);
    SFH_HIP(hipMemcpy(Z, z.data(), T_ * sizeof(double2), hipMemcpyHostToDevice));
    SFH_HIP(hipMemcpy(P, gp, T_ * sizeof(double), hipMemcpyHostToDevice));
}

void Solver::download_cotangent(const double2* Z, const double* P, double* gx, double* gy, double* gp) {
    std::vector<double2> z(T_);
    SFH_HIP(hipStreamSynchronize(stream_));
    SFH_HIP(hipMemcpy(z.data(), Z, T_ * sizeof(double2), hipMemcpyDeviceToHost));
    SFH_HIP(hipMemcpy(gp, P, T_ * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 0; i < T_; ++i) {
        gx[i] = z[i].x;
        gy[i] = z[i].y;
    }
}

void Solver::rhs_vjp(const double* gvx, const double* gvy, const double* gdphi, double* gx, double* gy,
                     double* gphi, const double* gdD, double* gD) {
    ensure_vjp();
    if (ph_.augmented != (gdD != nullptr && gD != nullptr))
        throw ArgError(ph_.augmented ? "the augmented drive's reverse pass takes the cotangents of D (sfh_rhs_vjp_augmented)"
                                     : "the handle has no augmented drive");
    ensure_vjp_d();
    SFH_HIP(hipStreamSynchronize(stream_));
    upload_cotangent(gvx, gvy, gdphi, gA_, gAP_);
    if (gdD) SFH_HIP(hipMemcpy(gAD_, gdD, T_ * sizeof(double), hipMemcpyHostToDevice));
    StageCtl c{};
    c.stage = 0;
    c.Z0 = Z0_;
    c.P0 = P0_;
    for (int i = 0; i < 3; ++i) {
        c.kZ[i] = vK_[i];
        c.kP[i] = vKP_[i];
    }
    set_delayed_ctl(c, D0_, nullptr);
    enqueue_rhs(Z0_, P0_, c);   // the forward mu at the current state
    SFH_HIP(hipMemcpyAsync(vMu_[0], mu_, T_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    enqueue_vjp(Z0_, P0_, vMu_[0], gA_, gAP_, 1.0, nullptr, nullptr, 0.0, gR_[0], gRP_[0], false, -1, gAD_, nullptr,
                gRD_[0]);
    download_cotangent(gR_[0], gRP_[0], gx, gy, gphi);
    if (gD) SFH_HIP(hipMemcpy(gD, gRD_[0], T_ * sizeof(double), hipMemcpyDeviceToHost));
    check_status();
}

void Solver::ensure_vjp_d() {
    if (!ph_.augmented || gAD_) return;
    gAD_ = dalloc<double>(T_);
    gAccD_ = dalloc<double>(T_);
    for (double** b : {&gRD_[0], &gRD_[1], &vYD_[0], &vYD_[1], &vYD_[2]}) *b = dalloc<double>(T_);
}

// Adjoint of one classical RK4 step from y_n (AutonomousRungeKuttaStepper.cuh:338-375):
//   k1 = f(y_n), y2 = y_n + dt/2 k1, k2 = f(y2), y3 = y_n + dt/2 k2, k3 = f(y3), y4 = y_n + dt k3,
//   k4 = f(y4), y_{n+1} = y_n + dt/6 (k1 + 2 k2 + 2 k3 + k4).
// With a = cotangent of y_{n+1} and r_s = J(y_s)^T kbar_s:
//   r4 from kbar4 = dt/6 a;  r3 from kbar3 = dt/3 a + dt r4;  r2 from kbar2 = dt/3 a + dt/2 r3;
//   r1 from kbar1 = dt/6 a + dt/2 r2;  cotangent of y_n = a + r1 + r2 + r3 + r4.
void Solver::step_vjp(double dt, const double* gx_in, const double* gy_in, const double* gp_in, double* gx_out,
                      double* gy_out, double* gp_out, const double* gD_in, double* gD_out) {
    ensure_vjp();
    if (ph_.augmented != (gD_in != nullptr && gD_out != nullptr))
        throw ArgError(ph_.augmented ? "the augmented drive's reverse step takes the cotangents of D (sfh_step_vjp_augmented)"
                                     : "the handle has no augmented drive");
    ensure_vjp_d();
    SFH_HIP(hipStreamSynchronize(stream_));
    upload_cotangent(gx_in, gy_in, gp_in, gA_, gAP_);
    if (gD_in) SFH_HIP(hipMemcpy(gAD_, gD_in, T_ * sizeof(double), hipMemcpyHostToDevice));
    enqueue_step_vjp(dt, Z0_, P0_, nullptr);
    download_cotangent(gAcc_, gAccP_, gx_out, gy_out, gp_out);
    if (gD_out) SFH_HIP(hipMemcpy(gD_out, gAccD_, T_ * sizeof(double), hipMemcpyDeviceToHost));
    check_status();
}

void Solver::enqueue_step_vjp(double dt, const double2* Z, const double* P, const StageCapture* stored) {
    SFH_HIP(hipMemcpyAsync(gAcc_, gA_, T_ * sizeof(double2), hipMemcpyDeviceToDevice, stream_));
    SFH_HIP(hipMemcpyAsync(gAccP_, gAP_, T_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    const bool aug = ph_.augmented;
    if (aug) SFH_HIP(hipMemcpyAsync(gAccD_, gAD_, T_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    StageCapture sc;
    if (stored) {
        sc = *stored;
    } else {   // stage states and their mu recomputed
        StageCtl c{};
        c.Z0 = Z;
        c.P0 = P;
        for (int i = 0; i < 3; ++i) {
            c.kZ[i] = vK_[i];
            c.kP[i] = vKP_[i];
        }
        const double2* zin[4] = {Z, vY_[0], vY_[1], vY_[2]};
        const double* pin[4] = {P, vYP_[0], vYP_[1], vYP_[2]};
        const double alpha[3] = {0.5 * dt, 0.5 * dt, dt};
        seq_len_ = gm_m_;   // standalone solves: the full GMRES launch sequence
        for (int st = 0; st < 4; ++st) {
            c.stage = st < 3 ? st + 1 : 0;
            c.alpha = st < 3 ? alpha[st] : 0.0;
            c.Zout = st < 3 ? vY_[st] : nullptr;
            c.Pout = st < 3 ? vYP_[st] : nullptr;
            // the augmented drive's D: stage input D0 or the previous stage's value, output the next
            set_delayed_ctl(c, st == 0 ? D0_ : vYD_[st - 1], st < 3 ? vYD_[st] : nullptr);
            enqueue_rhs(zin[st], pin[st], c);
            SFH_HIP(hipMemcpyAsync(vMu_[st], mu_, T_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
        }
        for (int i = 0; i < 3; ++i) {
            sc.Y[i] = vY_[i];
            sc.YP[i] = vYP_[i];
        }
        for (int i = 0; i < 4; ++i) sc.mu[i] = vMu_[i];
    }
    // reverse sweep
    const double* gAD = aug ? gAD_ : nullptr;
    enqueue_vjp(sc.Y[2], sc.YP[2], sc.mu[3], gA_, gAP_, dt / 6.0, nullptr, nullptr, 0.0, gR_[0], gRP_[0], true, 3, gAD,
                nullptr, aug ? gRD_[0] : nullptr);
    enqueue_vjp(sc.Y[1], sc.YP[1], sc.mu[2], gA_, gAP_, dt / 3.0, gR_[0], gRP_[0], dt, gR_[1], gRP_[1], true, 2, gAD,
                aug ? gRD_[0] : nullptr, aug ? gRD_[1] : nullptr);
    enqueue_vjp(sc.Y[0], sc.YP[0], sc.mu[1], gA_, gAP_, dt / 3.0, gR_[1], gRP_[1], 0.5 * dt, gR_[0], gRP_[0], true, 1,
                gAD, aug ? gRD_[1] : nullptr, aug ? gRD_[0] : nullptr);
    enqueue_vjp(Z, P, sc.mu[0], gA_, gAP_, dt / 6.0, gR_[0], gRP_[0], 0.5 * dt, gR_[1], gRP_[1], true, 0, gAD,
                aug ? gRD_[0] : nullptr, aug ? gRD_[1] : nullptr);
}

Solver::StageCapture Solver::seg_capture(int k) const {
    StageCapture c;
    const size_t T = T_, cap = seg_cap_;
    for (int i = 0; i < 3; ++i) {
        c.Y[i] = segY_ + (i * cap + k) * T;
        c.YP[i] = segYP_ + (i * cap + k) * T;
    }
    for (int i = 0; i < 4; ++i) c.mu[i] = segMu_ + (i * cap + k) * T;
    return c;
}

// The same resets as set_state (a new trajectory), from states already in device memory.
void Solver::load_state(const double2* Z, const double* P) {
    clear_solver_state();
    SFH_HIP(hipMemcpyAsync(Z0_, Z, T_ * sizeof(double2), hipMemcpyDeviceToDevice, stream_));
    SFH_HIP(hipMemcpyAsync(P0_, P, T_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    if (gmhvalid_) SFH_HIP(hipMemsetAsync(gmhvalid_, 0, 4 * sizeof(int), stream_));
    hist_steps_ = 0;
    time_ = drive_t_ = drive_prev_ = 0.0;
    if (D0_) SFH_HIP(hipMemsetAsync(D0_, 0, T_ * sizeof(double), stream_));
    if (bhvalid_) SFH_HIP(hipMemsetAsync(bhvalid_, 0, 4 * S_ * sizeof(int), stream_));
}

int Solver::traj_forward(double dt, int steps, int every) {
    if (steps < 1 || every < 1) throw ArgError("a trajectory needs steps >= 1 and checkpoint_every >= 1");
    ensure_vjp();   // the problem must have a reverse pass: rejected before any step
    if (ph_.augmented) throw ArgError("device trajectories take the two-field state; the augmented drive has three");
    const int nck = (steps + every - 1) / every;
    TrajRecord r;
    r.dt = dt;
    r.steps = steps;
    r.every = every;
    r.Z = dalloc<double2>((size_t)nck * T_);
    try {
        r.P = dalloc<double>((size_t)nck * T_);
    } catch (...) {
        (void)hipFree(r.Z);
        throw;
    }
    const int id = traj_next_++;
    traj_[id] = r;
    try {
        TrajRecord& rr = traj_[id];
        double* hist = opt_.solver == kSolveGMRES ? gmhist_ : opt_.solver == kSolveBatched ? bhist_ : nullptr;
        int* hvalid = opt_.solver == kSolveGMRES ? gmhvalid_ : opt_.solver == kSolveBatched ? bhvalid_ : nullptr;
        rr.hist_len = hist ? (size_t)4 * kHistDepth * T_ : 0;
        rr.hv_len = !hvalid ? 0 : opt_.solver == kSolveGMRES ? 4 : (size_t)4 * S_;
        rr.Mu = dalloc<double>((size_t)nck * T_);
        if (hist) {
            rr.H = dalloc<double>((size_t)nck * rr.hist_len);
            rr.HV = dalloc<int>((size_t)nck * rr.hv_len);
        }
        rr.hist_steps.assign(nck, 0);
        for (int c = 0; c < nck; ++c) {
            SFH_HIP(hipMemcpyAsync(rr.Z + (size_t)c * T_, Z0_, T_ * sizeof(double2), hipMemcpyDeviceToDevice, stream_));
            SFH_HIP(hipMemcpyAsync(rr.P + (size_t)c * T_, P0_, T_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
            SFH_HIP(hipMemcpyAsync(rr.Mu + (size_t)c * T_, mu_, T_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
            if (hist) {
                SFH_HIP(hipMemcpyAsync(rr.H + c * rr.hist_len, hist, rr.hist_len * sizeof(double),
                                       hipMemcpyDeviceToDevice, stream_));
                SFH_HIP(hipMemcpyAsync(rr.HV + c * rr.hv_len, hvalid, rr.hv_len * sizeof(int), hipMemcpyDeviceToDevice,
                                       stream_));
            }
            rr.hist_steps[c] = hist_steps_;
            step(dt, std::min(every, steps - c * every));
        }
        check_status();
    } catch (...) {
        traj_release(id);
        throw;
    }
    return id;
}

void Solver::traj_backward(int record, const double* gx_in, const double* gy_in, const double* gp_in,
                           double* gx_out, double* gy_out, double* gp_out) {
    auto it = traj_.find(record);
    if (it == traj_.end()) throw ArgError("unknown trajectory record " + std::to_string(record));
    const TrajRecord r = it->second;
    ensure_vjp();
    const int seg_len = std::min(r.every, r.steps);
    if (seg_len > seg_cap_) {
        SFH_HIP(hipStreamSynchronize(stream_));
        for (void* p : {(void*)segZ_, (void*)segP_, (void*)segY_, (void*)segYP_, (void*)segMu_})
            if (p) SFH_HIP(hipFree(p));
        segZ_ = nullptr;
        segP_ = nullptr;
        segY_ = nullptr;
        segYP_ = nullptr;
        segMu_ = nullptr;
        seg_cap_ = 0;
        segZ_ = dalloc<double2>((size_t)seg_len * T_);
        segP_ = dalloc<double>((size_t)seg_len * T_);
        segY_ = dalloc<double2>((size_t)3 * seg_len * T_);
        segYP_ = dalloc<double>((size_t)3 * seg_len * T_);
        segMu_ = dalloc<double>((size_t)4 * seg_len * T_);
        seg_cap_ = seg_len;
    }
    SFH_HIP(hipStreamSynchronize(stream_));
    upload_cotangent(gx_in, gy_in, gp_in, gA_, gAP_);
    const int nck = (r.steps + r.every - 1) / r.every;
    const size_t zb = T_ * sizeof(double2), pb = T_ * sizeof(double);
    // transposed solves (LU and GMRES handles): warm-started from each stage's solutions of the
    // previous reverse steps; the first reverse step with five of them for every stage measures the
    // Arnoldi steps its solves take, and from the next segment on the solves get that many launches
    // plus a spare.  A segment in which one ran out is swept again with full sequences from its saved
    // start, and the length doubles.
    const bool mwg = opt_.solver != kSolveBatched;
    struct SweepState {   // the warm-start state lives only inside this sweep, whatever ends it
        Solver* s;
        ~SweepState() {
            s->adj_warm_ = false;
            s->adj_calib_ = false;
            s->adj_len_ = 0;
        }
    } sweep_state{this};
    adj_warm_ = mwg;
    adj_len_ = 0;
    for (int st = 0; st < 4; ++st) adj_head_[st] = adj_hv_[st] = 0;
    bool calibrated = false;
    int pending_len = 0;
    for (int seg = nck - 1; seg >= 0; --seg) {
        const int n = std::min(r.every, r.steps - seg * r.every);
        // y_s, ..., y_{s+n-1}: the checkpoint, then n - 1 steps from it, each step keeping its stage
        // states and stage mu (the last step runs only for them).  The solves are warm-started from
        // the stage histories as in the forward pass; GMRES solves get the full launch sequence while
        // the histories fill and the forward pass's trimmed length after that, and a segment in which
        // a trimmed solve ran out of launches is run again with full sequences.
        const double2* ckZ = r.Z + (size_t)seg * T_;
        const double* ckP = r.P + (size_t)seg * T_;
        // the steps from the checkpoint before the stage histories hold kHistDepth solutions get the full
        // sequence (only in a trajectory's first segment: later checkpoints keep full histories)
        const int cold = std::max(0, kCalibSteps - r.hist_steps[seg]);
        auto rerun = [&](bool trimmed) {
            SFH_HIP(hipMemcpyAsync(segZ_, ckZ, zb, hipMemcpyDeviceToDevice, stream_));
            SFH_HIP(hipMemcpyAsync(segP_, ckP, pb, hipMemcpyDeviceToDevice, stream_));
            // the checkpoint's state and the solver's warm-start state there, as the forward pass had them
            load_state(ckZ, ckP);
            SFH_HIP(hipMemcpyAsync(mu_, r.Mu + (size_t)seg * T_, pb, hipMemcpyDeviceToDevice, stream_));
            if (r.hist_len) {
                double* hist = opt_.solver == kSolveGMRES ? gmhist_ : bhist_;
                int* hvalid = opt_.solver == kSolveGMRES ? gmhvalid_ : bhvalid_;
                SFH_HIP(hipMemcpyAsync(hist, r.H + seg * r.hist_len, r.hist_len * sizeof(double), hipMemcpyDeviceToDevice,
                                       stream_));
                SFH_HIP(hipMemcpyAsync(hvalid, r.HV + seg * r.hv_len, r.hv_len * sizeof(int), hipMemcpyDeviceToDevice,
                                       stream_));
            }
            hist_steps_ = r.hist_steps[seg];
            for (int k = 0; k < n; ++k) {
                seq_len_ = (trimmed && k >= cold) ? trim_ : gm_m_;
                const StageCapture cap = seg_capture(k);
                enqueue_step(r.dt, &cap);
                if (k + 1 < n) {
                    SFH_HIP(hipMemcpyAsync(segZ_ + (size_t)(k + 1) * T_, Z0_, zb, hipMemcpyDeviceToDevice, stream_));
                    SFH_HIP(hipMemcpyAsync(segP_ + (size_t)(k + 1) * T_, P0_, pb, hipMemcpyDeviceToDevice, stream_));
                }
            }
        };
        const bool trim = opt_.solver == kSolveGMRES && trim_ >= 2 && trim_ < gm_m_ && n > cold;   // calibrated
        if (trim) {
            sync();
            const int unc0 = unc_raw_;
            rerun(true);
            sync();
            if (unc_raw_ > unc0) {
                unc_discard_ += unc_raw_ - unc0;
                stats_.unconverged = unc_raw_ - unc_discard_;
                stats_.redone_steps += n;
                rerun(false);
            }
        } else {
            rerun(false);
        }
        auto sweep = [&]() {
            for (int i = n - 1; i >= 0; --i) {
                bool calib = mwg && !calibrated;
                for (int st = 0; st < 4; ++st) calib = calib && adj_hv_[st] >= kHistDepth;
                if (calib) {
                    adj_calib_ = true;
                    adj_calib_max_ = 0;
                }
                const StageCapture cap = seg_capture(i);
                enqueue_step_vjp(r.dt, segZ_ + (size_t)i * T_, segP_ + (size_t)i * T_, &cap);
                if (calib) {
                    adj_calib_ = false;
                    calibrated = true;
                    pending_len = std::min(gm_m_, adj_calib_max_ + 2);
                    if (opt_.force_trim > 0) pending_len = std::max(2, std::min(gm_m_, opt_.force_trim));
                }
                SFH_HIP(hipMemcpyAsync(gA_, gAcc_, zb, hipMemcpyDeviceToDevice, stream_));
                SFH_HIP(hipMemcpyAsync(gAP_, gAccP_, pb, hipMemcpyDeviceToDevice, stream_));
            }
        };
        const bool trimmed = mwg && adj_len_ > 0;
        if (trimmed) {
            const size_t ring = (size_t)4 * kHistDepth * T_ * sizeof(double);
            SFH_HIP(hipMemcpyAsync(adjSaveZ_, gA_, zb, hipMemcpyDeviceToDevice, stream_));
            SFH_HIP(hipMemcpyAsync(adjSaveP_, gAP_, pb, hipMemcpyDeviceToDevice, stream_));
            SFH_HIP(hipMemcpyAsync(adjSave_, adjRing_, ring, hipMemcpyDeviceToDevice, stream_));
            int head[4], hv[4];
            std::copy(adj_head_, adj_head_ + 4, head);
            std::copy(adj_hv_, adj_hv_ + 4, hv);
            sync();
            const int unc0 = unc_raw_;
            sweep();
            sync();
            if (unc_raw_ > unc0) {
                unc_discard_ += unc_raw_ - unc0;
                stats_.unconverged = unc_raw_ - unc_discard_;
                stats_.redone_steps += n;
                SFH_HIP(hipMemcpyAsync(gA_, adjSaveZ_, zb, hipMemcpyDeviceToDevice, stream_));
                SFH_HIP(hipMemcpyAsync(gAP_, adjSaveP_, pb, hipMemcpyDeviceToDevice, stream_));
                SFH_HIP(hipMemcpyAsync(adjRing_, adjSave_, ring, hipMemcpyDeviceToDevice, stream_));
                std::copy(head, head + 4, adj_head_);
                std::copy(hv, hv + 4, adj_hv_);
                const int len = adj_len_;
                adj_len_ = 0;
                sweep();
                adj_len_ = std::min(gm_m_, 2 * len);
            }
        } else {
            sweep();
            if (calibrated && pending_len > 0) adj_len_ = pending_len;   // trimmed from the next segment on
        }
    }
    load_state(r.Z, r.P);   // the record's initial state
    download_cotangent(gA_, gAP_, gx_out, gy_out, gp_out);
    check_status();
}

void Solver::free_record(TrajRecord& r) {
    for (void* p : {(void*)r.Z, (void*)r.P, (void*)r.Mu, (void*)r.H, (void*)r.HV})
        if (p) (void)hipFree(p);
    r.Z = nullptr;
    r.P = nullptr;
    r.Mu = nullptr;
    r.H = nullptr;
    r.HV = nullptr;
}

void Solver::traj_release(int record) {
    auto it = traj_.find(record);
    if (it == traj_.end()) throw ArgError("unknown trajectory record " + std::to_string(record));
    SFH_HIP(hipStreamSynchronize(stream_));
    free_record(it->second);
    traj_.erase(it);
}

}  // namespace sfh
