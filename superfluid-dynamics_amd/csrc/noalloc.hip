// noalloc.hip -- right-hand side of one surface without any device allocation.
//
// sfh_rhs_noalloc computes what sfh_rhs computes for a fresh SFH_SOLVE_BATCHED FP64 handle of one
// surface (spectral derivatives, dipole matrix, in-workgroup GMRES solve from a zero initial
// guess, mu', velocity and Bernoulli: batched_stage_kernel at stage 0), from device arrays to
// device arrays on the caller's stream.  Every buffer beyond the caller's inputs and outputs lives
// in `scratch`, a file-scope device array of kScratchBytes (one copy per GPU, held until the
// process ends):
//   M (n x n FP64, column-major), Z, Z', Z'', mu', velocity (n double2 each), b, mu (n doubles),
//   the twiddle table (n double2), the surface descriptor, the workgroup order and the stats.
// Host code never passes the array's name to a kernel: it takes the array's device address per
// GPU from hipGetSymbolAddress on the first call on that GPU and carves the buffers from it.
// Nothing is allocated: no hipMalloc, no hipFree, no library call that allocates.  The setup
// kernel clears what must start at zero (the initial guess mu and the stats) on every call,
// since the array keeps the previous call's contents.  Calls that overlap on one GPU race on the
// array.
#include <mutex>

#include "sfh_physics.hpp"
#include "solver.hpp"

namespace sfh {

constexpr int kNoAllocMaxN = kBatchMaxN;   // 1024 nodes
constexpr size_t kScratchAlign = 256;
constexpr size_t align_up(size_t v) { return (v + kScratchAlign - 1) / kScratchAlign * kScratchAlign; }
// the batched kernel's tile-skewed matrix: at most 1024^2 elements for every N <= 1024
constexpr size_t kScrM = align_up(sizeof(double) * tiled_matrix_elems(kNoAllocMaxN));
static_assert(tiled_matrix_elems(kNoAllocMaxN) == (long long)kNoAllocMaxN * kNoAllocMaxN, "full tiles at N = 1024");
constexpr size_t kScrVec2 = align_up(sizeof(double2) * kNoAllocMaxN);
constexpr size_t kScrVec = align_up(sizeof(double) * kNoAllocMaxN);
constexpr size_t kScrSmall = 256;   // descriptor, order, stats
// M, six double2 vectors (Z, Z', Z'', mu', velocity, twiddles), two double vectors (b, mu), small
constexpr size_t kScratchBytes = kScrM + 6 * kScrVec2 + 2 * kScrVec + kScrSmall;
static_assert(sizeof(SurfDesc) + 2 * 4 * sizeof(int) <= kScrSmall, "small block holds the descriptor");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct NoAllocBufs {
    double* M;
    double2 *Z, *Zp, *Zpp, *mup, *q, *tw;
    double *b, *mu;
    SurfDesc* desc;
    int *order, *stats;
};

static NoAllocBufs carve(unsigned char* base) {
    NoAllocBufs B;
    size_t o = 0;
    B.M = reinterpret_cast<double*>(base + o);
    o += kScrM;
    double2** v2[6] = {&B.Z, &B.Zp, &B.Zpp, &B.mup, &B.q, &B.tw};
    for (double2** p : v2) {
        *p = reinterpret_cast<double2*>(base + o);
        o += kScrVec2;
    }
    B.b = reinterpret_cast<double*>(base + o);
    o += kScrVec;
    B.mu = reinterpret_cast<double*>(base + o);
    o += kScrVec;
    B.desc = reinterpret_cast<SurfDesc*>(base + o);
    B.order = reinterpret_cast<int*>(base + o + sizeof(SurfDesc));
    B.stats = B.order + 4;
    return B;
}

// Descriptor, order, stats, zero initial guess, twiddles exp(-2 pi i t / n) and Z = x + i y.
__global__ __launch_bounds__(256) void noalloc_setup_kernel(NoAllocBufs B, SurfDesc d, const double* __restrict__ x,
                                                            const double* __restrict__ y) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) {
        *B.desc = d;
        B.order[0] = 0;
        for (int i = 0; i < 4; ++i) B.stats[i] = 0;
    }
    if (t < d.n) {
        B.mu[t] = 0.0;
        double s, c;
        sincospi(2.0 * (double)t / (double)d.n, &s, &c);
        B.tw[t] = make_double2(c, -s);
        B.Z[t] = make_double2(x[t], y[t]);
    }
}

__global__ __launch_bounds__(256) void noalloc_unpack_kernel(NoAllocBufs B, int n, double* __restrict__ vx,
                                                             double* __restrict__ vy, double* __restrict__ mu) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double2 q = B.q[t];
    vx[t] = q.x;
    vy[t] = q.y;
    if (mu) mu[t] = B.mu[t];
}

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// Device address of this GPU's scratch array, looked up once per device with that device current.
static unsigned char* scratch_address() {
    int dev = 0;
    SFH_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw HipError("device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        SFH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw HipError("hipGetSymbolAddress returned no address");
        g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
    }
    return g_scratch_ptr[dev];
}

size_t noalloc_scratch_bytes() { return kScratchBytes; }

void rhs_noalloc(const Phys& ph, int n, const double* x, const double* y, const double* phi, double* vx,
                 double* vy, double* dphi, double* mu, double tol, int max_iter, hipStream_t stream) {
    if (n < 1 || n > kNoAllocMaxN) throw ArgError("unsupported node count " + std::to_string(n) + ": the scratch array holds one matrix of at most 1024 nodes");
    int m = n;
    if (m % 3 == 0) m /= 3;
    if ((m & (m - 1)) != 0) throw ArgError("unsupported node count " + std::to_string(n) + " (n = 2^k or 3 * 2^k)");
    if (!x || !y || !phi || !vx || !vy || !dphi) throw ArgError("null array");
    if (!(tol > 0.0) || max_iter < 1) throw ArgError("tol must be positive and max_iter at least 1");
    const NoAllocBufs B = carve(scratch_address());
    SurfDesc d{};
    d.n = n;
    d.r3 = (n % 3 == 0) ? 1 : 0;
    int l = 0;
    while ((1 << l) < m) ++l;
    d.log2m = l;
    const int grid = (n + 255) / 256;
    hipLaunchKernelGGL(noalloc_setup_kernel, dim3(grid), dim3(256), 0, stream, B, d, x, y);
    SFH_HIP(hipGetLastError());
    BatchArgs a{};
    a.sd = B.desc;
    a.order = B.order;
    a.tw_all = B.tw;
    a.Zin = B.Z;
    a.Pin = phi;
    a.M = B.M;
    a.Zp = B.Zp;
    a.Zpp = B.Zpp;
    a.mup = B.mup;
    a.b = B.b;
    a.mu = B.mu;
    a.stats = B.stats;
    a.tol = tol;
    a.mixed = 0;
    a.cycles = (max_iter + 7) / 8;   // restart cycles of 8 basis vectors
    a.max_refine = 6;
    a.nsurf = 1;
    a.total = n;
    a.hist = nullptr;
    a.hvalid = nullptr;
    StageCtl c{};
    c.stage = 0;
    c.kZ[0] = B.q;
    c.kP[0] = dphi;
    c.Pin = phi;
    hipLaunchKernelGGL(batched_stage_kernel, dim3(1), dim3(kBatchThreads), 0, stream, a, ph, c);
    SFH_HIP(hipGetLastError());
    hipLaunchKernelGGL(noalloc_unpack_kernel, dim3(grid), dim3(256), 0, stream, B, n, vx, vy, mu);
    SFH_HIP(hipGetLastError());
}

}  // namespace sfh
