// fs_host.cpp -- host side of libfluidsim_hip.so: device buffers, multigrid hierarchy, the projection
// driver and the extern "C" ABI declared in include/fluidsim_hip.h.
//
// One fs_grid owns every buffer of one simulation on one device.  A projection is a fixed sequence of
// kernels on the context's stream; the PCG iteration body is captured once into a hipGraph and
// replayed.  All loop decisions (alpha, beta, convergence, restarts, iteration budget) live in a
// device scalar block, so the host only polls a status word between batches of iterations.
#include "fs_internal.hpp"
#include "fs_json.hpp"
#include "fs_kernels.hpp"
#include "../../include/fluidsim_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <limits>
#include <map>
#include <type_traits>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace fsim {

namespace {

struct StatusError : std::runtime_error {
    fs_status code;
    StatusError(fs_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void fail(fs_status code, const std::string& message) { throw StatusError(code, message); }

void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        fail(FS_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPC(call) hip_check((call), #call)

}  // namespace

// Serialises the calls that implicitly synchronise with every blocking stream (allocation, freeing,
// synchronous copies and fills on the null stream) against hipGraph capture on another thread's
// stream, so that grids may be created, captured and destroyed concurrently from several threads.
std::mutex& device_setup_mutex()
{
    static std::mutex m;
    return m;
}

namespace {

// Size classes of the device cache: multiples of 4 KiB up to 64 KiB, then four classes per power of two
// (at most 25% over the request).
size_t size_class(size_t bytes)
{
    if (bytes <= 65536)
        return (bytes + 4095) & ~(size_t)4095;
    size_t top = 1;
    while (top < bytes)
        top <<= 1;
    const size_t step = top / 8;   // classes top/2 + k top/8, k = 1..4
    return ((bytes + step - 1) / step) * step;
}

struct DeviceCache {
    std::mutex m;
    std::map<std::pair<int, size_t>, std::vector<void*>> free_blocks;   // (device, class) -> blocks
    size_t cached = 0;
    std::map<int, hipStream_t> setup_streams;
    std::mutex setup_m;   // one set-up copy or fill at a time per process
    std::map<size_t, std::vector<void*>> pinned_free_blocks;
};

DeviceCache& cache()
{
    static DeviceCache* c = new DeviceCache();   // outlives static destructors that may still release
    return *c;
}

// Cached bytes beyond which released blocks are freed for real.
constexpr size_t kCacheLimit = (size_t)64 << 30;

int current_device()
{
    int d = 0;
    HIPC(hipGetDevice(&d));
    return d;
}

hipStream_t setup_stream_locked(int device)
{
    DeviceCache& c = cache();
    auto it = c.setup_streams.find(device);
    if (it != c.setup_streams.end())
        return it->second;
    hipStream_t st = nullptr;
    HIPC(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    c.setup_streams[device] = st;
    return st;
}

}  // namespace

void* device_alloc(size_t bytes)
{
    if (bytes == 0)
        return nullptr;
    const int device = current_device();
    const size_t cls = size_class(bytes);
    DeviceCache& c = cache();
    {
        std::lock_guard<std::mutex> lock(c.m);
        auto it = c.free_blocks.find({device, cls});
        if (it != c.free_blocks.end() && !it->second.empty()) {
            void* p = it->second.back();
            it->second.pop_back();
            c.cached -= cls;
            return p;
        }
    }
    void* p = nullptr;
    std::lock_guard<std::mutex> lock(device_setup_mutex());
    HIPC(hipMalloc(&p, cls));
    return p;
}

void device_free(void* p, size_t bytes)
{
    if (!p)
        return;
    const size_t cls = size_class(bytes);
    DeviceCache& c = cache();
    {
        std::lock_guard<std::mutex> lock(c.m);
        if (c.cached + cls <= kCacheLimit) {
            int device = 0;
            if (hipGetDevice(&device) == hipSuccess) {
                c.free_blocks[{device, cls}].push_back(p);
                c.cached += cls;
                return;
            }
        }
    }
    std::lock_guard<std::mutex> lock(device_setup_mutex());
    (void)hipFree(p);
}

void device_zero(void* p, size_t bytes)
{
    if (!p || bytes == 0)
        return;
    DeviceCache& c = cache();
    std::lock_guard<std::mutex> lock(c.setup_m);
    const hipStream_t st = setup_stream_locked(current_device());
    HIPC(hipMemsetAsync(p, 0, bytes, st));
    HIPC(hipStreamSynchronize(st));
}

void device_upload(void* dst, const void* src, size_t bytes)
{
    if (bytes == 0)
        return;
    DeviceCache& c = cache();
    std::lock_guard<std::mutex> lock(c.setup_m);
    const hipStream_t st = setup_stream_locked(current_device());
    HIPC(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
    HIPC(hipStreamSynchronize(st));
}

void* pinned_alloc(size_t bytes)
{
    DeviceCache& c = cache();
    {
        std::lock_guard<std::mutex> lock(c.m);
        auto it = c.pinned_free_blocks.find(bytes);
        if (it != c.pinned_free_blocks.end() && !it->second.empty()) {
            void* p = it->second.back();
            it->second.pop_back();
            return p;
        }
    }
    void* p = nullptr;
    std::lock_guard<std::mutex> lock(device_setup_mutex());
    HIPC(hipHostMalloc(&p, bytes, hipHostMallocDefault));
    return p;
}

void pinned_free(void* p, size_t bytes)
{
    if (!p)
        return;
    DeviceCache& c = cache();
    std::lock_guard<std::mutex> lock(c.m);
    c.pinned_free_blocks[bytes].push_back(p);
}

namespace {

template <class T> struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void swap(DevBuf& o) noexcept
    {
        std::swap(p, o.p);
        std::swap(n, o.n);
    }
    // Returns the memory to the device cache: the owner has drained every stream that used it.
    void release()
    {
        device_free(p, n * sizeof(T));
        p = nullptr;
        n = 0;
    }
    // Allocates and zero-fills (halos of padded fields must start, and stay, zero), releasing any previous
    // allocation.  Returns the change in allocated bytes (negative when a larger buffer is replaced), which
    // callers add to their byte count (unsigned wrap-around gives the right total).
    long long alloc(size_t count)
    {
        const long long before = (long long)(n * sizeof(T));
        release();
        n = count;
        if (count == 0)
            return -before;
        p = (T*)device_alloc(count * sizeof(T));
        device_zero(p, count * sizeof(T));
        return (long long)(count * sizeof(T)) - before;
    }
};

struct Level {
    int w = 0, h = 0, s = 0;
    std::vector<double> cx, cy;
    // transfers to the next coarser level (empty on the last level)
    std::vector<int> pxf, pxs, pyf, pys;
    std::vector<double> pxw, pyw;
    std::vector<int> rxc, rxi, ryc, ryi;
    std::vector<double> rxw, ryw;

    DevBuf<double> d_cx, d_cy, right64, down64;
    DevBuf<float> right32, down32, invd32, rhs, sol0, sol1, res;
    DevBuf<int> d_pxf, d_pxs, d_pyf, d_pys, d_rxc, d_rxi, d_ryc, d_ryi;
    DevBuf<float> d_pxw, d_pyw, d_rxw, d_ryw;
    // 16-bit copies (fs_set_multigrid_precision): conductances of a tiled level, and the bfloat16 rhs and
    // solutions of levels 1 .. tail start
    DevBuf<unsigned short> right16, down16, rhs16, sol16a, sol16b;

    int smooth_variant = 0;     // tile of k_rb_multi pre-smoothing on this level (autotuned)
    int post_variant = 0;       // tile of k_rb_multi prolongation + post-smoothing (autotuned)
    int restrict_variant = 0;   // tile of k_res_restrict from this level (autotuned)
    int jacobi_pre = 0;         // tiles of k_jacobi_multi pre- and post-smoothing (autotuned on first use)
    int jacobi_post = 0;
    int jacobi_fused = -1;      // k_jacobi_res_restrict tile for pre-smoothing + restriction; -1: two launches

    size_t padded() const { return (size_t)s * (h + 2); }
};

// buildTransfers for one axis (PressureMultigrid.cpp:64-98): bilinear prolongation on cell centres and
// restriction as its exact transpose.
void build_axis(const std::vector<double>& fc, const std::vector<double>& cc, std::vector<int>& first,
    std::vector<int>& second, std::vector<double>& weight, std::vector<int>& count, std::vector<int>& index,
    std::vector<double>& rweight)
{
    const int nf = (int)fc.size(), nc = (int)cc.size();
    first.assign(nf, 0);
    second.assign(nf, 0);
    weight.assign(nf, 0.0);
    count.assign(nc, 0);
    index.assign((size_t)nc * 4, -1);
    rweight.assign((size_t)nc * 4, 0.0);
    for (int f = 0; f < nf; ++f) {
        const int parent = f / 2;
        const int a = std::max(0, parent - int(fc[f] < cc[parent]));
        const int b = std::min(a + 1, nc - 1);
        const double wgt = a == b ? 0.0 : std::clamp((fc[f] - cc[a]) / (cc[b] - cc[a]), 0.0, 1.0);
        first[f] = a;
        second[f] = b;
        weight[f] = wgt;
        auto append = [&](int c, double value) {
            if (value == 0.0)
                return;
            if (count[c] >= 4)
                fail(FS_E_RUNTIME, "multigrid restriction stencil exceeds four taps");
            const int k = count[c]++;
            index[(size_t)c * 4 + k] = f;
            rweight[(size_t)c * 4 + k] = value;
        };
        append(a, 1.0 - wgt);
        append(b, wgt);
    }
}

template <class T> void upload_vec(DevBuf<T>& d, const std::vector<T>& h, size_t& bytes)
{
    bytes += d.alloc(h.size());
    device_upload(d.p, h.data(), h.size() * sizeof(T));
}

std::vector<float> to_float(const std::vector<double>& v)
{
    std::vector<float> r(v.size());
    for (size_t i = 0; i < v.size(); ++i)
        r[i] = (float)v[i];
    return r;
}

thread_local std::string g_last_error;

}  // namespace

void set_last_error(const std::string& message) { g_last_error = message; }

std::string solid_mask_problem(const unsigned char* mask, int W, int H)
{
    const size_t N = (size_t)W * H;
    size_t fluid = 0, start = N;
    for (size_t i = 0; i < N; ++i)
        if (!mask[i]) {
            ++fluid;
            if (start == N)
                start = i;
        }
    if (fluid == 0)
        return "a solid mask must leave at least one fluid cell";
    std::vector<unsigned char> seen(N, 0);
    std::vector<size_t> stack{start};
    seen[start] = 1;
    size_t reached = 0;
    while (!stack.empty()) {
        const size_t i = stack.back();
        stack.pop_back();
        ++reached;
        const int x = (int)(i % W), y = (int)(i / W);
        auto visit = [&](size_t j) {
            if (!mask[j] && !seen[j]) {
                seen[j] = 1;
                stack.push_back(j);
            }
        };
        if (x > 0) visit(i - 1);
        if (x + 1 < W) visit(i + 1);
        if (y > 0) visit(i - W);
        if (y + 1 < H) visit(i + W);
    }
    return reached == fluid ? std::string() : "the fluid cells of a solid mask must form one 4-connected region";
}

}  // namespace fsim

using namespace fsim;

struct fs_context {
    int device = 0;
    size_t lds_per_block = 0;
    int compute_units = 0;
    hipStream_t stream = nullptr;
    hipStream_t aux = nullptr;          // second stream (default priority): hierarchy rebuild overlap
    hipEvent_t fork = nullptr, join = nullptr;
    hipEvent_t join_levels = nullptr;   // the rebuild's tiled levels are done (the coarse chain may still run)
    PcgScalars* pinned = nullptr;   // host-pinned mirror of a grid's scalar block
    double* pinned_value = nullptr; // host-pinned landing slot of single device reductions
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    DevBuf<BatchInstance> batch_desc;       // descriptors of the last batched solve
    DevBuf<unsigned char> vis_batch_desc;   // descriptors of the last batched viscosity solve
    PcgScalars* batch_pinned = nullptr;     // host-pinned landing slots of a batch's scalar blocks
    size_t batch_pinned_n = 0;
    unsigned char* staging = nullptr;       // host-pinned source of the batch descriptors' upload
    size_t staging_n = 0;
    long long batch_solo_cells = 0;         // fs_context_set_batch_split (0: every instance batched)
    hipStream_t batch_stream = nullptr;     // the persistent launch's stream while solo solves run
    hipEvent_t batch_in = nullptr, batch_out = nullptr;
};

constexpr int kPressureBudget = 100000;   // PCG iterations of an unbounded (negative) budget
constexpr size_t kFoldBlocks = 1536;      // cell blocks (64 x 16) beyond which PCG partials are pre-reduced
// Multigrid levels above this many cells take the grid's chosen smoother (fs_set_pressure_smoother); smaller
// ones, the single-workgroup tail's levels among them, always run red-black V(4,4).
constexpr long long kSmootherMinCells = 4096;
// Levels of at most this many cells are coarsened by the hierarchy rebuild's single-workgroup launch.
constexpr long long kChainCells = 4096;

// Weight i of `sweeps` Chebyshev-weighted Jacobi sweeps: PressureMultigridGPU's constructor
// (PressureMultigridGPU.cpp:24-28), 1 / (1.1 + 0.9 cos(pi (i + 1/2) / sweeps)) -- the reciprocals of the
// Chebyshev nodes of [0.2, 2], the spectrum it assumes for D^-1 A -- rounded to float.
float chebyshev_weight(int i, int sweeps)
{
    const double pi = 3.14159265358979323846;
    return (float)(1.0 / (1.1 + 0.9 * std::cos(pi * (i + 0.5) / sweeps)));
}
constexpr size_t kBatchSpreadLds = 84 * 1024;   // LDS of a batched-solve workgroup: one per compute unit

// A batched solve split by fs_context_set_batch_split: grids above the split take their single-grid solve
// on the context's stream while the persistent launch runs on a stream of its own.
bool batch_solo(const fs_context* ctx, int W, int H)
{
    return ctx->aux && ctx->batch_solo_cells > 0 && (long long)W * H > ctx->batch_solo_cells;
}

// The persistent launch's stream of a split batch, ordered after the work already on the context's stream.
hipStream_t batch_side_stream(fs_context* ctx)
{
    if (!ctx->batch_stream) {
        std::lock_guard<std::mutex> lock(device_setup_mutex());
        HIPC(hipStreamCreate(&ctx->batch_stream));
        HIPC(hipEventCreateWithFlags(&ctx->batch_in, hipEventDisableTiming));
        HIPC(hipEventCreateWithFlags(&ctx->batch_out, hipEventDisableTiming));
    }
    HIPC(hipEventRecord(ctx->batch_in, ctx->stream));
    HIPC(hipStreamWaitEvent(ctx->batch_stream, ctx->batch_in, 0));
    return ctx->batch_stream;
}

// Waits for a stream when it goes out of scope, also when an exception leaves the scope: host memory
// that an enqueued copy reads must outlive the copy.
struct SyncOnExit {
    hipStream_t stream;
    ~SyncOnExit() { (void)hipStreamSynchronize(stream); }
};

// Copies `bytes` of host memory to device memory on the context's stream through the context's pinned
// staging buffer, so the copy neither depends on the caller's memory staying alive nor reads pageable
// memory after this returns.  The caller synchronises the stream before the staging buffer is reused
// (every batched entry point does before it returns).
void stage_upload(fs_context* ctx, void* dst, const void* src, size_t bytes)
{
    if (bytes == 0)
        return;
    if (ctx->staging_n < bytes) {
        HIPC(hipStreamSynchronize(ctx->stream));
        std::lock_guard<std::mutex> lock(device_setup_mutex());
        if (ctx->staging)
            (void)hipHostFree(ctx->staging);
        ctx->staging = nullptr;
        ctx->staging_n = 0;
        HIPC(hipHostMalloc((void**)&ctx->staging, bytes, hipHostMallocDefault));
        ctx->staging_n = bytes;
    }
    std::memcpy(ctx->staging, src, bytes);
    HIPC(hipMemcpyAsync(dst, ctx->staging, bytes, hipMemcpyHostToDevice, ctx->stream));
}

struct fs_grid {
    fs_context* ctx = nullptr;
    int W = 0, H = 0;
    double h = 1.0;
    double ox = 0.0, oy = 0.0;
    double gravity_y = 100.0;  // FluidGridCPU.h:37
    double sigma = 0.0;
    std::vector<double> rho_l, rho_g, mu_l, mu_g;
    fs_preconditioner precond = FS_PRECOND_MULTIGRID;
    fs_transport transport = FS_TRANSPORT_REFERENCE;
    fs_mg_precision mg_prec = FS_MG_FP32;
    bool flexible_cg = false;   // fs_set_flexible_cg
    bool relative_tol = false;  // fs_set_relative_tolerance
    fs_pressure_smoother smoother = FS_SMOOTHER_RED_BLACK;   // fs_set_pressure_smoother: levels above
    int smooth_sweeps = 4;                                    // kSmootherMinCells cells
    int coarse_sweeps = 4;                                    // red-black sweeps of the smaller levels
    bool jacobi_tuned = false;
    bool uploaded = false;
    bool coeff_valid = false;
    bool mg_valid = false;
    bool has_snapshot = false;
    size_t bytes = 0;
    int predicted = 8;

    DevBuf<double> fraction, u, v, lmass, gmass, density, p, p_prev, b, tmp64;
    DevBuf<double> snap_fraction, snap_u, snap_v, snap_p, snap_p_prev, snap_lmass, snap_gmass;
    DevBuf<double> sl_u, sl_v, sl_cells;
    // reference transport: second fraction/mass buffers, transported velocities (ping-pong), momenta,
    // per-pass face transfers, frozen corrections
    DevBuf<double> tr_fraction, tr_lmass, tr_gmass, tr_nu[2], tr_nv[2], tr_mom_u, tr_mom_v, tr_lt, tr_mt;
    DevBuf<double> tr_lcw, tr_gcw, tr_mcw, tr_out;
    DevBuf<unsigned char> tr_split;
    DevBuf<float> calib_src, calib_dst;   // 64 MiB copy-bandwidth reference (allocated on first use)
    ViscositySolver* visc = nullptr;       // implicit viscosity workspace (created on first use)
    int vis_sweeps = 0;                    // fs_set_viscosity_smoother (0: the reference's V(2,2))
    DevBuf<double2> st_heights, st_orient, st_raw, st_curv;   // capillary geometry (first use)
    DevBuf<float> x, r, d, q, zdiag;
    DevBuf<unsigned char> solid;   // solid-mask extension (empty: no mask)
    DevBuf<float> color_px;        // writeColorPixels output, 4 floats per cell (first use)
    DevBuf<EditCell> edit_cells;   // resolved material edits of the last commit (first use)
    long long solid_cells = 0;
    DevBuf<PcgScalars> S;
    DevBuf<PcgScalars> S_aux;   // failure flag of a hierarchy rebuilt on the second stream
    DevBuf<double> partials;
    DevBuf<double> factor;
    DevBuf<float> ginv;
    int coarse_n = 0;
    size_t tail_lds = 0;
    std::vector<Level> levels;
    int tail_start = 0;
    TailDesc tail{};
    DevBuf<TailDesc> tail_dev;
    DevBuf<unsigned> tail_tables;

    hipGraphExec_t iter_exec = nullptr;
    hipGraph_t iter_graph = nullptr;
    fs_preconditioner graph_precond = FS_PRECOND_MULTIGRID;

    ~fs_grid()
    {
        drop_graph();
        viscosity_destroy(visc);
    }

    // capillaryStepLimit (FluidGridCPU.cpp:843-860)
    double capillary_step_limit() const
    {
        if (sigma == 0.0 || rho_l.empty() || rho_g.empty())
            return std::numeric_limits<double>::infinity();
        double ml = rho_l.front(), mg = rho_g.front();
        for (double d : rho_l)
            ml = std::min(ml, d);
        for (double d : rho_g)
            mg = std::min(mg, d);
        constexpr double pi = 3.14159265358979323846;
        return 0.5 * std::sqrt((ml + mg) * h * h * h / (2.0 * pi * sigma));
    }

    // applySurfaceTension (FluidGridCPU.cpp:862-926): geometry, then the face force with the FP64
    // pressure face weights of the current density.
    void apply_surface_tension(double dt)
    {
        const size_t n = N();
        if (st_curv.n != n) {
            bytes += st_heights.alloc(n);
            bytes += st_orient.alloc(n);
            bytes += st_raw.alloc(n);
            bytes += st_curv.alloc(n);
        }
        launch_capillary_geometry(st(), W, H, h, fraction.p, st_heights.p, st_orient.p, st_raw.p, st_curv.p);
        rebuild_coefficients();
        const Level& L0 = levels[0];
        launch_capillary_force(st(), W, H, h, fraction.p, st_curv.p, L0.right64.p, L0.down64.p, sigma, dt, u.p, v.p);
    }

    bool has_viscosity() const
    {
        for (double mu : mu_l)
            if (mu > 0.0)
                return true;
        for (double mu : mu_g)
            if (mu > 0.0)
                return true;
        return false;
    }

    size_t N() const { return (size_t)W * H; }
    size_t NP() const { return (size_t)(W + 2) * (H + 2); }
    size_t NU() const { return (size_t)(W + 1) * H; }
    size_t NV() const { return (size_t)W * (H + 1); }
    hipStream_t st() const { return ctx->stream; }

    // Waits for every stream of the context that may still use this grid's buffers, before any of them is
    // released or replaced (released memory goes straight back to the device cache).
    void drain() const
    {
        HIPC(hipStreamSynchronize(ctx->stream));
        if (ctx->aux)
            HIPC(hipStreamSynchronize(ctx->aux));
        if (ctx->batch_stream)
            HIPC(hipStreamSynchronize(ctx->batch_stream));
    }

    void drop_graph()
    {
        if (iter_exec)
            (void)hipGraphExecDestroy(iter_exec);
        if (iter_graph)
            (void)hipGraphDestroy(iter_graph);
        iter_exec = nullptr;
        iter_graph = nullptr;
    }

    GridDev dev()
    {
        GridDev g{};
        g.W = W;
        g.H = H;
        g.s = W + 2;
        g.h = h;
        g.nl = (int)rho_l.size();
        g.ng = (int)rho_g.size();
        g.fraction = fraction.p;
        g.u = u.p;
        g.v = v.p;
        g.lmass = lmass.p;
        g.gmass = gmass.p;
        g.density = density.p;
        g.p = p.p;
        g.p_prev = p_prev.p;
        g.b = b.p;
        g.right64 = levels[0].right64.p;
        g.down64 = levels[0].down64.p;
        g.right32 = levels[0].right32.p;
        g.down32 = levels[0].down32.p;
        g.invd32 = levels[0].invd32.p;
        g.x = x.p;
        g.r = r.p;
        g.d = d.p;
        g.q = q.p;
        g.zdiag = zdiag.p;
        g.S = S.p;
        g.partials = partials.p;
        const size_t blocks = (size_t)((W + 63) / 64) * ((H + 15) / 16);
        g.partials_ad = partials.p + 5 * blocks;
        // every consumer workgroup reading every block's partial costs blocks^2 loads: past this many
        // blocks (a 1024 x 1536 grid) one extra single-workgroup launch pre-reduces them
        const bool fold = blocks > kFoldBlocks;
        g.fold_ad = fold ? partials.p + 6 * blocks : g.partials_ad;
        g.fold_ad_n = fold ? 1 : (int)blocks;
        g.fold_rz = fold ? partials.p + 6 * blocks + 8 : partials.p;
        g.fold_rz_n = fold ? 1 : (int)blocks;
        g.solid = solid.n ? solid.p : nullptr;
        return g;
    }

    LevelView view(int l)
    {
        Level& L = levels[l];
        LevelView v{};
        v.w = L.w;
        v.h = L.h;
        v.s = L.s;
        v.right = L.right32.p;
        v.down = L.down32.p;
        v.invd = L.invd32.p;
        v.rhs = l == 0 ? r.p : L.rhs.p;
        v.sol[0] = L.sol0.p;
        v.sol[1] = L.sol1.p;
        v.res = L.res.p;
        return v;
    }

    const int* status_ptr() const { return &S.p->status; }
    const double* rmean_ptr() const { return &S.p->r_mean; }

    // ---------------------------------------------------------------------------------------
    void allocate()
    {
        const size_t NPd = NP();
        bytes += fraction.alloc(N());
        bytes += u.alloc(NU());
        bytes += v.alloc(NV());
        bytes += density.alloc(N());
        bytes += p.alloc(NPd);
        bytes += p_prev.alloc(NPd);
        bytes += b.alloc(NPd);
        bytes += tmp64.alloc(N());
        bytes += x.alloc(NPd);
        bytes += r.alloc(NPd);
        bytes += d.alloc(NPd);
        bytes += q.alloc(NPd);
        bytes += zdiag.alloc(NPd);
        bytes += S.alloc(1);
        bytes += S_aux.alloc(1);
        const size_t blocks = (size_t)((W + 63) / 64) * ((H + 15) / 16);
        // region A: up to five partials per cell block; region B (after it): the operator's d.q partials,
        // read by the update kernel while it writes region A; then 16 slots of pre-reduced sums
        bytes += partials.alloc(6 * blocks + 64);
        build_hierarchy();
        tune_tiles();
        if (smoother == FS_SMOOTHER_CHEBYSHEV)
            tune_jacobi();
    }

    void fill_tail(int start)
    {
        std::memset(&tail, 0, sizeof(tail));
        const int nlev = (int)levels.size();
        tail.sweeps = coarse_sweeps;
        tail.levels = nlev - start;
        tail.coarse_n = coarse_n;
        tail.ginv = ginv.p;
        tail.top_rhs = start == 0 ? r.p : levels[start].rhs.p;
        tail.top_sol = levels[start].sol0.p;
        for (int i = 0; i < tail.levels; ++i) {
            Level& L = levels[start + i];
            TailLevel& T = tail.lev[i];
            T.w = L.w;
            T.h = L.h;
            T.s = L.s;
            const bool last = start + i + 1 == nlev;
            T.cw = last ? 0 : levels[start + i + 1].w;
            T.ch = last ? 0 : levels[start + i + 1].h;
            T.right = L.right32.p;
            T.down = L.down32.p;
            T.pxf = L.d_pxf.p; T.pxs = L.d_pxs.p; T.pxw = L.d_pxw.p;
            T.pyf = L.d_pyf.p; T.pys = L.d_pys.p; T.pyw = L.d_pyw.p;
            T.rxc = L.d_rxc.p; T.rxi = L.d_rxi.p; T.rxw = L.d_rxw.p;
            T.ryc = L.d_ryc.p; T.ryi = L.d_ryi.p; T.ryw = L.d_ryw.p;
        }
    }

    // Every tail level's transfer tables in the kernel's LDS layout, as one word block.
    void build_tail_tables()
    {
        std::vector<unsigned> words((size_t)std::max(tail.table_words, 1), 0u);
        auto fbits = [](double v) {
            const float f = (float)v;
            unsigned u;
            std::memcpy(&u, &f, 4);
            return u;
        };
        for (int i = 0; i < tail.levels; ++i) {
            const TailLevel& T = tail.lev[i];
            const Level& L = levels[tail_start + i];
            size_t o = (size_t)(T.lds_tables - tail.lds_tables_base);
            auto put_i = [&](const std::vector<int>& v, size_t n) {
                for (size_t k = 0; k < n; ++k)
                    words[o + k] = k < v.size() ? (unsigned)v[k] : 0u;
                o += n;
            };
            auto put_f = [&](const std::vector<double>& v, size_t n) {
                for (size_t k = 0; k < n; ++k)
                    words[o + k] = k < v.size() ? fbits(v[k]) : 0u;
                o += n;
            };
            put_i(L.pxf, T.w);
            put_i(L.pxs, T.w);
            put_f(L.pxw, T.w);
            put_i(L.pyf, T.h);
            put_i(L.pys, T.h);
            put_f(L.pyw, T.h);
            put_i(L.rxc, T.cw);
            put_i(L.rxi, 4 * (size_t)T.cw);
            put_f(L.rxw, 4 * (size_t)T.cw);
            put_i(L.ryc, T.ch);
            put_i(L.ryi, 4 * (size_t)T.ch);
            put_f(L.ryw, 4 * (size_t)T.ch);
        }
        upload_vec(tail_tables, words, bytes);
        tail.tables = tail_tables.p;
    }

    // PressureMultigrid constructor (PressureMultigrid.cpp:38-62): halve until w*h <= 64.
    void build_hierarchy()
    {
        levels.clear();
        int w = W, hh = H, spacing = 1;
        for (;;) {
            levels.emplace_back();
            Level& L = levels.back();
            L.w = w;
            L.h = hh;
            L.s = w + 2;
            L.cx.resize(w);
            L.cy.resize(hh);
            for (int i = 0; i < w; ++i)
                L.cx[i] = 0.5 * (i * spacing + std::min((i + 1) * spacing, W));
            for (int i = 0; i < hh; ++i)
                L.cy[i] = 0.5 * (i * spacing + std::min((i + 1) * spacing, H));
            if (w * hh <= 64)
                break;
            w = (w + 1) / 2;
            hh = (hh + 1) / 2;
            spacing *= 2;
        }
        for (size_t l = 0; l + 1 < levels.size(); ++l) {
            Level& F = levels[l];
            Level& Cl = levels[l + 1];
            build_axis(F.cx, Cl.cx, F.pxf, F.pxs, F.pxw, F.rxc, F.rxi, F.rxw);
            build_axis(F.cy, Cl.cy, F.pyf, F.pys, F.pyw, F.ryc, F.ryi, F.ryw);
        }
        for (size_t l = 0; l < levels.size(); ++l) {
            Level& L = levels[l];
            const size_t np = L.padded();
            upload_vec(L.d_cx, L.cx, bytes);
            upload_vec(L.d_cy, L.cy, bytes);
            bytes += L.right64.alloc(np);
            bytes += L.down64.alloc(np);
            bytes += L.right32.alloc(np);
            bytes += L.down32.alloc(np);
            bytes += L.invd32.alloc(np);
            if (l > 0)
                bytes += L.rhs.alloc(np);
            bytes += L.sol0.alloc(np);
            bytes += L.sol1.alloc(np);
            bytes += L.res.alloc(np);
            if (l + 1 < levels.size()) {
                upload_vec(L.d_pxf, L.pxf, bytes);
                upload_vec(L.d_pxs, L.pxs, bytes);
                upload_vec(L.d_pyf, L.pyf, bytes);
                upload_vec(L.d_pys, L.pys, bytes);
                upload_vec(L.d_rxc, L.rxc, bytes);
                upload_vec(L.d_rxi, L.rxi, bytes);
                upload_vec(L.d_ryc, L.ryc, bytes);
                upload_vec(L.d_ryi, L.ryi, bytes);
                upload_vec(L.d_pxw, to_float(L.pxw), bytes);
                upload_vec(L.d_pyw, to_float(L.pyw), bytes);
                upload_vec(L.d_rxw, to_float(L.rxw), bytes);
                upload_vec(L.d_ryw, to_float(L.ryw), bytes);
            }
        }
        const Level& last = levels.back();
        coarse_n = last.w * last.h - 1;
        bytes += factor.alloc((size_t)std::max(coarse_n, 1) * std::max(coarse_n, 1));

        bytes += ginv.alloc((size_t)std::max(coarse_n, 1) * std::max(coarse_n, 1));

        // Coarse tail: the deepest levels whose LDS footprint fits one workgroup.
        const int nlev = (int)levels.size();
        size_t limit = ctx->lds_per_block > 0 ? ctx->lds_per_block : 65536;
        limit = std::min<size_t>(limit, 160 * 1024) - 4096;
        for (;;) {
            tail_start = nlev - 1;
            for (int l = nlev - 2; l >= 0 && nlev - l <= kMaxTailLevels; --l) {
                fill_tail(l);
                size_t elements = 0;
                for (int i = 0; i < tail.levels; ++i)
                    elements += (size_t)tail.lev[i].s * (tail.lev[i].h + 2);
                // the tail kernel stages at most 8 padded elements and 2 table words per thread and
                // gives each of its 1024 threads at most 4 cells (2 coarse cells) of a level
                const Level& L = levels[l];
                const Level& Cn = levels[l + 1];
                const bool fits_threads = (long)L.w * L.h <= 4096 && (long)((L.w + 1) / 2) * L.h <= 4096 &&
                                          (long)Cn.w * Cn.h <= 2048;
                if (tail_layout(tail) > limit || elements > 8 * 1024 || tail.table_words > 2 * 1024 || !fits_threads)
                    break;
                tail_start = l;
            }
            fill_tail(tail_start);
            tail_lds = tail_layout(tail);
            build_tail_tables();
            if (tail_lds <= 65536 || tail_allow_lds(tail_lds) || limit <= 65536)
                break;
            limit = 65536 - 4096;  // the device refused a larger dynamic LDS request
        }
        bytes += tail_dev.alloc(1);
        device_upload(tail_dev.p, &tail, sizeof(TailDesc));
    }

    // ---------------------------------------------------------------------------------------
    void rebuild_coefficients()
    {
        GridDev g = dev();
        launch_density(st(), g);
        launch_face_weights(st(), g);
        Level& L0 = levels[0];
        launch_level_diag(st(), W, H, L0.right64.p, L0.down64.p, L0.right32.p, L0.down32.p, L0.invd32.p);
        coeff_valid = true;
        mg_valid = false;
    }

    // PressureMultigrid::rebuild (PressureMultigrid.cpp:119-140) on the device, FP64, on `stream`; a
    // failed coarse factorisation is flagged in `flag`.
    // The levels of more than kChainCells cells are built one launch pair each; the smaller ones (at most
    // kMaxTailLevels of them) and the coarse inverse in one single-workgroup launch, whose phases are
    // too small to fill the device -- the same values either way.
    // `levels_done` (optional) is recorded on `stream` once the levels the tiled V-cycle kernels read are
    // built and packed; returns true when the coarse chain (tail levels and coarse inverse) is launched
    // after it, i.e. only the tail launch has to wait for the rest.
    bool rebuild_hierarchy(hipStream_t stream, PcgScalars* flag, hipEvent_t levels_done = nullptr)
    {
        const size_t nlev = levels.size();
        size_t chain_from = nlev;   // first level built by the chain launch
        while (chain_from > 1 && nlev - (chain_from - 1) <= (size_t)kMaxTailLevels &&
               (long long)levels[chain_from - 1].w * levels[chain_from - 1].h <= kChainCells)
            --chain_from;
        for (size_t l = 0; l + 1 < chain_from; ++l) {
            Level& F = levels[l];
            Level& Cl = levels[l + 1];
            launch_coarsen(stream, F.w, F.h, F.d_cx.p, F.d_cy.p, F.right64.p, F.down64.p, Cl.w, Cl.h, Cl.d_cx.p,
                Cl.d_cy.p, Cl.right64.p, Cl.down64.p);
            launch_level_diag(stream, Cl.w, Cl.h, Cl.right64.p, Cl.down64.p, Cl.right32.p, Cl.down32.p,
                Cl.invd32.p);
        }
        const Level& last = levels.back();
        // the chain builds tail levels only: pack the tiled levels and let the V-cycle start
        const bool early = levels_done && chain_from < nlev && (int)chain_from >= tail_start;
        if (early) {
            pack_levels(stream);
            HIPC(hipEventRecord(levels_done, stream));
        }
        if (chain_from < nlev) {
            ChainDesc chain{};
            for (size_t l = chain_from; l < nlev; ++l) {
                const Level& F = levels[l - 1];
                Level& Cl = levels[l];
                chain.lev[chain.levels++] = ChainLevel{F.w, F.h, Cl.w, Cl.h, F.d_cx.p, F.d_cy.p, F.right64.p,
                    F.down64.p, Cl.d_cx.p, Cl.d_cy.p, Cl.right64.p, Cl.down64.p, Cl.right32.p, Cl.down32.p,
                    Cl.invd32.p};
            }
            launch_coarse_chain(stream, chain, ginv.p, coarse_n, flag, solid.n > 0);
        } else {
            launch_coarse_inverse(stream, last.w, last.h, last.right64.p, last.down64.p, ginv.p, coarse_n, flag,
                solid.n > 0);
        }
        if (!early)
            pack_levels(stream);
        mg_valid = true;
        return early;
    }
    void rebuild_hierarchy() { rebuild_hierarchy(st(), S.p); }

    // ---- 16-bit multigrid levels ----
    bool active16() const { return mg_prec != FS_MG_FP32 && tail_start >= 1; }
    // The batched solve runs red-black sweeps (any count) on every level, not the Chebyshev smoother.
    bool batch_cycle() const
    {
        return smoother == FS_SMOOTHER_RED_BLACK || (long long)W * H <= kSmootherMinCells;
    }
    // Applies a smoother choice (the caller has drained the grid's streams).
    void set_smoother(fs_pressure_smoother kind, int sweeps, int coarse)
    {
        smoother = kind;
        smooth_sweeps = sweeps;
        if (coarse != coarse_sweeps) {
            coarse_sweeps = coarse;
            tail.sweeps = coarse;
            if (tail_dev.p)
                device_upload(tail_dev.p, &tail, sizeof(TailDesc));
        }
        drop_graph();
        if (kind == FS_SMOOTHER_CHEBYSHEV && !levels.empty() && !jacobi_tuned)
            tune_jacobi();
    }

    // Allocates (or releases) the 16-bit arrays of the current mode and points the tail's top level at
    // the storage its neighbouring tiled level uses.
    void apply_precision()
    {
        const bool on = active16();
        const bool vec16 = on && mg_prec == FS_MG_BF16_VECTORS;
        for (int l = 0; l < (int)levels.size(); ++l) {
            Level& L = levels[l];
            const size_t np = L.padded();
            auto want = [&](DevBuf<unsigned short>& b, bool need16) {
                if (need16 && b.n != np)
                    bytes += b.alloc(np);
                else if (!need16 && b.n) {
                    bytes -= b.n * sizeof(unsigned short);
                    b.release();
                }
            };
            want(L.right16, on && l < tail_start);
            want(L.down16, on && l < tail_start);
            const bool v = vec16 && l >= 1 && l <= tail_start;
            want(L.rhs16, v);
            want(L.sol16a, v);
            want(L.sol16b, v && l < tail_start);
        }
        if (vec16) {
            tail.top_rhs = levels[tail_start].rhs16.p;
            tail.top_sol = levels[tail_start].sol16a.p;
        } else {
            tail.top_rhs = tail_start == 0 ? r.p : levels[tail_start].rhs.p;
            tail.top_sol = levels[tail_start].sol0.p;
        }
        mg_valid = false;   // the 16-bit conductances are packed by the next hierarchy rebuild
        drop_graph();
    }

    void pack_levels(hipStream_t stream)
    {
        for (int l = 0; l < tail_start && active16(); ++l) {
            Level& L = levels[l];
            if (mg_prec == FS_MG_FP16)
                launch_level_pack(stream, L.w, L.h, L.right64.p, L.down64.p, (half16*)L.right16.p, (half16*)L.down16.p);
            else
                launch_level_pack(stream, L.w, L.h, L.right64.p, L.down64.p, (bf16*)L.right16.p, (bf16*)L.down16.p);
        }
    }

    // Level l with C conductances and V vectors: FP32 vectors, or the bfloat16 ones of levels >= 1.
    template <class C, class V> LevelViewT<C, V> view_t(int l)
    {
        Level& L = levels[l];
        LevelViewT<C, V> v{};
        v.w = L.w;
        v.h = L.h;
        v.s = L.s;
        v.right = (const C*)L.right16.p;
        v.down = (const C*)L.down16.p;
        v.invd = L.invd32.p;
        if constexpr (std::is_same<V, float>::value) {
            v.rhs = l == 0 ? r.p : L.rhs.p;
            v.sol[0] = L.sol0.p;
            v.sol[1] = L.sol1.p;
        } else {
            v.rhs = (V*)L.rhs16.p;
            v.sol[0] = (V*)L.sol16a.p;
            v.sol[1] = (V*)L.sol16b.p;
        }
        v.res = nullptr;
        return v;
    }

    template <class VC> CoarseLinkT<VC> link_t(int l)
    {
        const Level& F = levels[l];
        const Level& Cl = levels[l + 1];
        const VC* csol = std::is_same<VC, float>::value ? (const VC*)Cl.sol0.p : (const VC*)Cl.sol16a.p;
        return CoarseLinkT<VC>{csol, Cl.s, F.d_pxf.p, F.d_pxs.p, F.d_pxw.p, F.d_pyf.p, F.d_pys.p, F.d_pyw.p};
    }

    // Pre-smoothing (from_zero) or prolongation + post-smoothing of tiled level l with 16-bit storage;
    // VL is the vector type of levels >= 1.
    template <class C, class VL> void smooth16(int l, int cur, int first_color, int from_zero, const int* status,
        int variant)
    {
        const CoarseLinkT<VL> link = from_zero ? CoarseLinkT<VL>{} : link_t<VL>(l);
        if (l == 0)
            launch_rb_multi_t<C, float, VL>(st(), view_t<C, float>(0), cur, r.p, rmean_ptr(), first_color, from_zero,
                link, status, variant, sweeps_on(l));
        else {
            const LevelViewT<C, VL> L = view_t<C, VL>(l);
            launch_rb_multi_t<C, VL, VL>(st(), L, cur, L.rhs, nullptr, first_color, from_zero, link, status, variant,
                sweeps_on(l));
        }
    }

    template <class C, class VL> void restrict16(int l, int cur, const int* status, int variant)
    {
        const Level& F = levels[l];
        const AxisRestrict rx{F.d_rxc.p, F.d_rxi.p, F.d_rxw.p};
        const AxisRestrict ry{F.d_ryc.p, F.d_ryi.p, F.d_ryw.p};
        if (l == 0)
            launch_res_restrict_t<C, float, VL>(st(), view_t<C, float>(0), cur, r.p, rmean_ptr(), view_t<C, VL>(1), rx,
                ry, status, variant);
        else {
            const LevelViewT<C, VL> L = view_t<C, VL>(l);
            launch_res_restrict_t<C, VL, VL>(st(), L, cur, L.rhs, nullptr, view_t<C, VL>(l + 1), rx, ry, status,
                variant);
        }
    }

    template <class C, class VL> void vcycle16(const int* status)
    {
        const int T = tail_start;
        for (int l = 0; l < T; ++l) {
            smooth16<C, VL>(l, 0, 0, 1, status, levels[l].smooth_variant);
            restrict16<C, VL>(l, 1, status, levels[l].restrict_variant);
        }
        join_before_tail();
        launch_mg_tail(st(), tail, tail_lds, nullptr, status, std::is_same<VL, bf16>::value);
        for (int l = T - 1; l >= 0; --l)
            smooth16<C, VL>(l, 1, 1, 0, status, levels[l].post_variant);
    }

    template <class C, class VL> void tune_tiles16()
    {
        auto time = [&](auto&& launch) {
            launch();
            HIPC(hipEventRecord(ctx->ev0, st()));
            for (int k = 0; k < 3; ++k)
                launch();
            HIPC(hipEventRecord(ctx->ev1, st()));
            HIPC(hipEventSynchronize(ctx->ev1));
            float ms = 0.f;
            HIPC(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
            return ms;
        };
        for (int l = 0; l < tail_start; ++l) {
            Level& F = levels[l];
            auto pick = [&](int count, auto&& launch) {
                float best = std::numeric_limits<float>::infinity();
                int choice = 0;
                for (int v = 0; v < count; ++v) {
                    const float ms = time([&] { launch(v); });
                    if (ms < best) {
                        best = ms;
                        choice = v;
                    }
                }
                return choice;
            };
            F.smooth_variant = pick(kSmoothVariants, [&](int v) { smooth16<C, VL>(l, 0, 0, 1, nullptr, v); });
            F.post_variant = pick(kSmoothVariants, [&](int v) { smooth16<C, VL>(l, 1, 1, 0, nullptr, v); });
            F.restrict_variant = pick(kRestrictVariants, [&](int v) { restrict16<C, VL>(l, 0, nullptr, v); });
        }
        drop_graph();
    }

    void tune_for_precision()
    {
        if (!active16())
            tune_tiles();
        else if (mg_prec == FS_MG_FP16)
            tune_tiles16<half16, float>();
        else if (mg_prec == FS_MG_BF16)
            tune_tiles16<bf16, float>();
        else
            tune_tiles16<bf16, bf16>();
    }

    // The finest post-smoothing launch of a V-cycle in the grid's storage mode (kernel timing).
    void fine_post_smooth(int cur)
    {
        if (mg_prec == FS_MG_FP16 && active16())
            smooth16<half16, float>(0, cur, 1, 0, nullptr, levels[0].post_variant);
        else if (mg_prec == FS_MG_BF16 && active16())
            smooth16<bf16, float>(0, cur, 1, 0, nullptr, levels[0].post_variant);
        else if (active16())
            smooth16<bf16, bf16>(0, cur, 1, 0, nullptr, levels[0].post_variant);
        else if (tail_start > 0)
            smooth_level(0, cur, r.p, rmean_ptr(), false, link_t<float>(0), nullptr, -1);
        else
            launch_rb_multi(st(), view(0), cur, r.p, rmean_ptr(), 1, 0, CoarseLink{}, nullptr, 1, levels[0].post_variant);
    }

    // One V(4,4) cycle on the residual r (minus its mean) into levels[0].sol0.
    void vcycle(const int* status)
    {
        if (active16()) {
            if (mg_prec == FS_MG_FP16)
                vcycle16<half16, float>(status);
            else if (mg_prec == FS_MG_BF16)
                vcycle16<bf16, float>(status);
            else
                vcycle16<bf16, bf16>(status);
            return;
        }
        const int T = tail_start;
        const CoarseLink none{};
        for (int l = 0; l < T; ++l) {
            LevelView L = view(l);
            const float* rhs = l == 0 ? r.p : L.rhs;
            const double* shift = l == 0 ? rmean_ptr() : nullptr;
            const Level& F = levels[l];
            AxisRestrict rx{F.d_rxc.p, F.d_rxi.p, F.d_rxw.p};
            AxisRestrict ry{F.d_ryc.p, F.d_ryi.p, F.d_ryw.p};
            if (chebyshev_on(l) && F.jacobi_fused >= 0) {
                launch_jacobi_res_restrict(st(), L, 0, rhs, shift, chebyshev_weights(true), smooth_sweeps, view(l + 1),
                    rx, ry, status, F.jacobi_fused);
                continue;
            }
            smooth_level(l, 0, rhs, shift, true, none, status, -1);
            launch_res_restrict(st(), L, 1, rhs, shift, view(l + 1), rx, ry, status, F.restrict_variant);
        }
        join_before_tail();
        launch_mg_tail(st(), tail, tail_lds, T == 0 ? rmean_ptr() : nullptr, status);
        for (int l = T - 1; l >= 0; --l) {
            LevelView L = view(l);
            const float* rhs = l == 0 ? r.p : L.rhs;
            const double* shift = l == 0 ? rmean_ptr() : nullptr;
            smooth_level(l, 1, rhs, shift, false, link_t<float>(l), status, -1);
        }
    }

    // Smoother of tiled level l: the chosen one above kSmootherMinCells cells, red-black V(4,4) below.
    bool chebyshev_on(int l) const
    {
        return smoother == FS_SMOOTHER_CHEBYSHEV && (long long)levels[l].w * levels[l].h > kSmootherMinCells;
    }
    int sweeps_on(int l) const
    {
        return (long long)levels[l].w * levels[l].h > kSmootherMinCells ? smooth_sweeps : coarse_sweeps;
    }
    // Weights of the Chebyshev sweeps in the order a pre- (or, reversed, a post-) smoothing applies them.
    SmoothWeights chebyshev_weights(bool pre) const
    {
        SmoothWeights wt{};
        for (int k = 0; k < smooth_sweeps; ++k)
            wt.w[pre ? k : smooth_sweeps - 1 - k] = chebyshev_weight(k, smooth_sweeps);
        return wt;
    }
    // Pre-smoothing (from zero: reads nothing, writes sol[cur ^ 1]) or prolongation + post-smoothing of
    // tiled level l in FP32 storage; variant -1 takes the level's tuned tile.
    void smooth_level(int l, int cur, const float* rhs, const double* shift, bool pre, const CoarseLink& link,
        const int* status, int variant)
    {
        const Level& F = levels[l];
        if (chebyshev_on(l)) {
            launch_jacobi_multi(st(), view(l), cur, rhs, shift, pre ? 1 : 0, chebyshev_weights(pre), smooth_sweeps,
                link, status, variant >= 0 ? variant : pre ? F.jacobi_pre : F.jacobi_post);
        } else {
            launch_rb_multi(st(), view(l), cur, rhs, shift, pre ? 0 : 1, pre ? 1 : 0, link, status, l == 0,
                variant >= 0 ? variant : pre ? F.smooth_variant : F.post_variant, sweeps_on(l));
        }
    }

    // Autotuner of the Chebyshev smoother's tiles (first use): every variant on each level's buffers.
    void tune_jacobi()
    {
        auto time = [&](auto&& launch) {
            launch();
            HIPC(hipEventRecord(ctx->ev0, st()));
            for (int k = 0; k < 3; ++k)
                launch();
            HIPC(hipEventRecord(ctx->ev1, st()));
            HIPC(hipEventSynchronize(ctx->ev1));
            float ms = 0.f;
            HIPC(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
            return ms;
        };
        const CoarseLink none{};
        const SmoothWeights pre = chebyshev_weights(true), post = chebyshev_weights(false);
        for (int l = 0; l < tail_start; ++l) {
            Level& F = levels[l];
            if ((long long)F.w * F.h <= kSmootherMinCells)
                continue;
            const LevelView L = view(l);
            const float* rhs = l == 0 ? r.p : L.rhs;
            const CoarseLink link = link_t<float>(l);
            float best = std::numeric_limits<float>::infinity();
            for (int v = 0; v < kJacobiVariants; ++v) {
                const float ms = time([&] { launch_jacobi_multi(st(), L, 0, rhs, nullptr, 1, pre, smooth_sweeps, none, nullptr, v); });
                if (ms < best) {
                    best = ms;
                    F.jacobi_pre = v;
                }
            }
            best = std::numeric_limits<float>::infinity();
            for (int v = 0; v < kJacobiVariants; ++v) {
                const float ms = time([&] { launch_jacobi_multi(st(), L, 1, rhs, nullptr, 0, post, smooth_sweeps, link, nullptr, v); });
                if (ms < best) {
                    best = ms;
                    F.jacobi_post = v;
                }
            }
            // pre-smoothing + residual + restriction: the two launches against the fused one
            F.jacobi_fused = -1;
            const Level& Cn = levels[l + 1];
            if (2 * (long long)Cn.w >= F.w && 2 * (long long)Cn.h >= F.h) {
                const LevelView C = view(l + 1);
                const AxisRestrict rx{F.d_rxc.p, F.d_rxi.p, F.d_rxw.p};
                const AxisRestrict ry{F.d_ryc.p, F.d_ryi.p, F.d_ryw.p};
                best = time([&] {
                    launch_jacobi_multi(st(), L, 0, rhs, nullptr, 1, pre, smooth_sweeps, none, nullptr, F.jacobi_pre);
                    launch_res_restrict(st(), L, 1, rhs, nullptr, C, rx, ry, nullptr, F.restrict_variant);
                });
                for (int v = 0; v < kJacobiRestrictVariants; ++v) {
                    const float ms = time([&] {
                        launch_jacobi_res_restrict(st(), L, 0, rhs, nullptr, pre, smooth_sweeps, C, rx, ry, nullptr, v);
                    });
                    if (ms < best) {
                        best = ms;
                        F.jacobi_fused = v;
                    }
                }
            }
        }
        jacobi_tuned = true;
        drop_graph();
    }

    // Autotuner: times every tile variant of the level smoother (plain sweeps) and of the residual +
    // restriction on each level's own buffers and keeps the fastest.  The fused kernels compute the same
    // values with any tile, so the choice affects time only.
    void tune_tiles()
    {
        const CoarseLink none{};
        auto time = [&](auto&& launch) {
            launch();
            HIPC(hipEventRecord(ctx->ev0, st()));
            for (int k = 0; k < 3; ++k)
                launch();
            HIPC(hipEventRecord(ctx->ev1, st()));
            HIPC(hipEventSynchronize(ctx->ev1));
            float ms = 0.f;
            HIPC(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
            return ms;
        };
        for (int l = 0; l < tail_start; ++l) {
            Level& F = levels[l];
            const LevelView L = view(l);
            const float* rhs = l == 0 ? r.p : L.rhs;
            const Level& Cl = levels[l + 1];
            const CoarseLink link{Cl.sol0.p, Cl.s, F.d_pxf.p, F.d_pxs.p, F.d_pxw.p, F.d_pyf.p, F.d_pys.p, F.d_pyw.p};
            float best = std::numeric_limits<float>::infinity();
            for (int v = 0; v < kSmoothVariants; ++v) {
                const float ms = time([&] { launch_rb_multi(st(), L, 0, rhs, nullptr, 0, 1, none, nullptr, l == 0, v); });
                if (ms < best) {
                    best = ms;
                    F.smooth_variant = v;
                }
            }
            best = std::numeric_limits<float>::infinity();
            for (int v = 0; v < kSmoothVariants; ++v) {
                const float ms = time([&] { launch_rb_multi(st(), L, 1, rhs, nullptr, 1, 0, link, nullptr, l == 0, v); });
                if (ms < best) {
                    best = ms;
                    F.post_variant = v;
                }
            }
            const AxisRestrict rx{F.d_rxc.p, F.d_rxi.p, F.d_rxw.p};
            const AxisRestrict ry{F.d_ryc.p, F.d_ryi.p, F.d_ryw.p};
            const LevelView C = view(l + 1);
            best = std::numeric_limits<float>::infinity();
            for (int v = 0; v < kRestrictVariants; ++v) {
                const float ms = time([&] { launch_res_restrict(st(), L, 0, rhs, nullptr, C, rx, ry, nullptr, v); });
                if (ms < best) {
                    best = ms;
                    F.restrict_variant = v;
                }
            }
        }
        drop_graph();   // launches captured before the choice keep the old tiles
    }

    const float* z_ptr() const { return precond == FS_PRECOND_MULTIGRID ? levels[0].sol0.p : zdiag.p; }
    // bfloat16 level vectors round on every store, so the V-cycle is linear only to ~2^-9: the CG then
    // takes the flexible (Polak-Ribiere) beta, which tolerates a varying preconditioner
    int flexible() const
    {
        return flexible_cg || (precond == FS_PRECOND_MULTIGRID && mg_prec == FS_MG_BF16_VECTORS && active16());
    }

    // A rebuild whose coarse chain still runs on the helper stream: the first tail launch waits for it
    // and takes over its failure flag (the kernels before it read only the tiled levels).
    hipEvent_t tail_join = nullptr;
    void join_before_tail()
    {
        if (!tail_join)
            return;
        HIPC(hipStreamWaitEvent(st(), tail_join, 0));
        launch_merge_failure(st(), S_aux.p, S.p);
        tail_join = nullptr;
    }

    void precondition(const int* status)
    {
        if (precond == FS_PRECOND_MULTIGRID)
            vcycle(status);
        else
            launch_precond_diag(st(), dev());
    }

    // The loop body (FluidGridCPU.cpp:1779-1819), every decision taken on the device.
    void iteration_kernels()
    {
        GridDev g = dev();
        launch_pcg_ad(st(), g);
        launch_pcg_update(st(), g);
        launch_flush(st(), g);
        launch_true_residual(st(), g, 1);
        precondition(status_ptr());
        launch_pcg_rz(st(), g, z_ptr(), flexible());
        launch_pcg_direction(st(), g, z_ptr(), flexible());
    }

    void ensure_graph()
    {
        if (iter_exec && graph_precond == precond)
            return;
        drop_graph();
        std::lock_guard<std::mutex> lock(device_setup_mutex());
        HIPC(hipStreamBeginCapture(st(), hipStreamCaptureModeRelaxed));
        iteration_kernels();
        HIPC(hipStreamEndCapture(st(), &iter_graph));
        HIPC(hipGraphInstantiate(&iter_exec, iter_graph, nullptr, nullptr, 0));
        graph_precond = precond;
    }

    PcgScalars read_scalars()
    {
        HIPC(hipMemcpyAsync(ctx->pinned, S.p, sizeof(PcgScalars), hipMemcpyDeviceToHost, st()));
        HIPC(hipStreamSynchronize(st()));
        return *ctx->pinned;
    }

    // FluidGridCPU::project (FluidGridCPU.cpp:1030-1124), first phase: scalar block, coefficients,
    // hierarchy, rhs, predicted pressure, its residual and the first direction -- launches only.
    void project_begin(double dt, int max_iterations, double tol)
    {
        const double scale = h * h / dt;
        if (relative_tol)
            launch_init_scalars(st(), S.p, 0.0, max_iterations, tol);   // tol * max|b|, set on the device
        else
            launch_init_scalars(st(), S.p, tol * scale, max_iterations);
        if (!coeff_valid)
            rebuild_coefficients();
        GridDev g = dev();
        // The hierarchy depends only on the face weights: rebuild it on the second stream while the main
        // stream forms the rhs, the predicted pressure and its residual; join before the first V-cycle.
        const bool rebuild = precond == FS_PRECOND_MULTIGRID && !mg_valid;
        // A context without the helper stream rebuilds in line.
        hipStream_t side = ctx->aux ? ctx->aux : st();
        bool late_tail = false;   // the first V-cycle's tail waits for the coarse chain
        if (rebuild) {
            if (ctx->aux) {
                HIPC(hipEventRecord(ctx->fork, st()));
                HIPC(hipStreamWaitEvent(side, ctx->fork, 0));
            }
            launch_init_scalars(side, S_aux.p, 0.0, 0);
            late_tail = rebuild_hierarchy(side, S_aux.p, ctx->aux ? ctx->join_levels : nullptr);
            if (ctx->aux)
                HIPC(hipEventRecord(ctx->join, side));
        }
        launch_rhs(st(), g, scale);
        launch_rhs_center(st(), g);
        launch_zero_pressure(st(), g);
        launch_predict(st(), g);
        launch_sub_mean_p(st(), g, 0);
        launch_true_residual(st(), g, 0);
        if (rebuild) {
            if (late_tail) {
                HIPC(hipStreamWaitEvent(st(), ctx->join_levels, 0));
                tail_join = ctx->join;   // consumed by the tail launch of the V-cycle below
            } else {
                if (ctx->aux)
                    HIPC(hipStreamWaitEvent(st(), ctx->join, 0));
                launch_merge_failure(st(), S_aux.p, S.p);
            }
        }
        {
            // a pending join never outlives this V-cycle, not even when a launch in it throws (a later
            // graph capture would record it)
            struct ClearJoin {
                hipEvent_t& e;
                ~ClearJoin() { e = nullptr; }
            } clear{tail_join};
            precondition(status_ptr());
        }
        launch_pcg_rz(st(), g, z_ptr(), flexible());
        launch_pcg_direction(st(), g, z_ptr(), flexible());
    }

    // Last phase: fold the correction into the pressure, remove its mean, correct the velocity and
    // measure the divergence -- launches only.  Every kernel of it checks on the device that the solve
    // has stopped and is a no-op otherwise, so it may be queued behind iterations that might not finish
    // the solve.
    void project_end(double dt)
    {
        GridDev g = dev();
        launch_final_flush_sum(st(), g);
        launch_sub_mean_p(st(), g, 1);
        launch_correct_velocity(st(), g, dt);
        launch_max_divergence(st(), g, true);
    }

    // `tol` is the caller's divergence tolerance; in the relative mode the accepted divergence is the
    // residual tolerance the device derived from max|b|, in divergence units (tol * max|b| * dt / h^2).
    fs_projection_result project_result(const PcgScalars& s, double tol, double dt)
    {
        if (relative_tol)
            tol = s.tol / (h * h / dt);
        fs_projection_result res{};
        res.max_divergence = std::numeric_limits<double>::infinity();
        res.physical_steps = 1;
        if (s.mg_failed)
            mg_valid = false;
        if (s.status == kFailed)
            return res;
        res.iterations = s.iterations;
        res.max_divergence = s.div_max;
        res.converged = s.div_max <= tol;
        res.restarts = s.restarts;
        res.initial_residual = s.initial_rmax;
        res.final_residual = s.true_rmax;
        res.rhs_max = s.bmax;
        return res;
    }

    // A negative budget runs to tolerance, the uniform WebGPU contract used by its scenes
    // (FluidGridGPU.cpp:1326, FluidScene.inl:355-359), bounded so that a stagnating solve ends.
    static int budget(int max_iterations) { return max_iterations < 0 ? kPressureBudget : max_iterations; }

    // FluidGridCPU::project (FluidGridCPU.cpp:1030-1124)
    fs_projection_result project(double dt, int max_iterations, double tol)
    {
        fs_projection_result res{};
        res.max_divergence = std::numeric_limits<double>::infinity();
        res.physical_steps = 1;
        if (!uploaded)
            fail(FS_E_LOGIC, "Upload the grid state before stepping or projecting");
        if (dt <= 0.0)
            return res;
        ensure_graph();
        project_begin(dt, budget(max_iterations), tol);
        // The set-up, the iterations the previous solve needed and the closing phase are queued without a
        // read-back in between (iterations past the solve's end and a closing phase before it are
        // device-side no-ops); more iterations follow, two at a time, while the solve still runs.
        int batch = std::min(predicted, 256);
        PcgScalars s{};
        for (;;) {
            for (int i = 0; i < batch; ++i)
                HIPC(hipGraphLaunch(iter_exec, st()));
            project_end(dt);
            s = read_scalars();
            if (s.status != kRunning)
                break;
            batch = 2;
        }
        predicted = s.iterations;
        return project_result(s, tol, dt);
    }

    // This grid's entry of a batched solve (the iteration body of iteration_kernels()).
    BatchInstance batch_instance()
    {
        if (tail_start > kMaxBatchLevels)
            fail(FS_E_UNSUPPORTED, "grid has more multigrid levels than a batched solve supports");
        BatchInstance b{};
        b.g = dev();
        b.multigrid = precond == FS_PRECOND_MULTIGRID;
        b.levels = tail_start;
        b.tail = tail_dev.p;
        for (int l = 0; l < tail_start; ++l) {
            const Level& F = levels[l];
            const Level& Cl = levels[l + 1];
            BatchLevel& B = b.lev[l];
            B.L = view(l);
            B.C = view(l + 1);
            B.rx = AxisRestrict{F.d_rxc.p, F.d_rxi.p, F.d_rxw.p};
            B.ry = AxisRestrict{F.d_ryc.p, F.d_ryi.p, F.d_ryw.p};
            B.link = CoarseLink{Cl.sol0.p, Cl.s, F.d_pxf.p, F.d_pxs.p, F.d_pxw.p, F.d_pyf.p, F.d_pys.p, F.d_pyw.p};
            B.sweeps = sweeps_on(l);
        }
        return b;
    }

    double max_divergence()
    {
        launch_max_divergence(st(), dev());
        return read_scalars().div_max;
    }
};

// ---------------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------------

namespace {

template <class F> fs_status guard(F&& f)
{
    try {
        f();
        g_last_error.clear();
        return FS_OK;
    } catch (const StatusError& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "out of host memory";
        return FS_E_RUNTIME;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return FS_E_RUNTIME;
    }
}

void need(bool cond, fs_status code, const char* message)
{
    if (!cond)
        fail(code, message);
}

void set_device(fs_context* ctx) { HIPC(hipSetDevice(ctx->device)); }

// ---- resident-instance checkpoint (FluidGridGPU.cpp:566-695 save, 697-880 load) ----
//
// Metadata: one JSON object naming the format, the grid geometry, every physical and solver setting
// and the solver history the next step depends on.  Payload: FP64 values in native byte order, in
// the fixed order of kCheckpointFields.  Derived data (density, face weights, the multigrid
// hierarchy, solver workspaces) is rebuilt by the next operation, as in the reference.

constexpr const char* kCheckpointBackend = "hip-uniform";
constexpr int kCheckpointVersion = 1;
const char* const kCheckpointFields[] = {"liquid_fraction", "u", "v", "pressure", "previous_pressure",
    "liquid_mass", "gas_mass"};
constexpr int kCheckpointFieldCount = 7;

const char* native_byte_order()
{
    const unsigned short probe = 1;
    unsigned char first = 0;
    std::memcpy(&first, &probe, 1);
    return first == 1 ? "little" : "big";
}

size_t checkpoint_values(const fs_grid& g, size_t n_liquid, size_t n_gas)
{
    return (3 + n_liquid + n_gas) * g.N() + g.NU() + g.NV();
}

std::string checkpoint_metadata(const fs_grid& g)
{
    using namespace fsjson;
    std::string out = "{";
    bool first = true;
    put_key(out, "backend", first);
    put_string(out, kCheckpointBackend);
    put_key(out, "version", first);
    put_int(out, kCheckpointVersion);
    put_key(out, "scalar", first);
    put_string(out, "float64");
    put_key(out, "byte_order", first);
    put_string(out, native_byte_order());
    put_key(out, "width", first);
    put_int(out, g.W);
    put_key(out, "height", first);
    put_int(out, g.H);
    put_key(out, "cell_size", first);
    put_number(out, g.h);
    put_key(out, "origin", first);
    out += '[';
    put_number(out, g.ox);
    out += ',';
    put_number(out, g.oy);
    out += ']';
    put_key(out, "gravity_y", first);
    put_number(out, g.gravity_y);
    put_key(out, "surface_tension", first);
    put_number(out, g.sigma);
    put_key(out, "preconditioner", first);
    put_int(out, (int)g.precond);
    put_key(out, "transport", first);
    put_int(out, (int)g.transport);
    auto mixture = [&](const char* key, const std::vector<double>& rho, const std::vector<double>& mu) {
        put_key(out, key, first);
        out += '[';
        for (size_t k = 0; k < rho.size(); ++k) {
            if (k)
                out += ',';
            bool inner = true;
            out += '{';
            put_key(out, "density", inner);
            put_number(out, rho[k]);
            // zero viscosity is omitted, as the reference does (FluidGridGPU.cpp:597-599)
            if (mu[k] != 0.0) {
                put_key(out, "dynamic_viscosity", inner);
                put_number(out, mu[k]);
            }
            out += '}';
        }
        out += ']';
    };
    mixture("liquid_mixture", g.rho_l, g.mu_l);
    mixture("gas_mixture", g.rho_g, g.mu_g);
    put_key(out, "previous_solve_iterations", first);
    put_int(out, g.predicted);
    put_key(out, "payload", first);
    out += '[';
    for (int i = 0; i < kCheckpointFieldCount; ++i) {
        if (i)
            out += ',';
        put_string(out, kCheckpointFields[i]);
    }
    out += "]}";
    return out;
}

[[noreturn]] void bad_checkpoint(const std::string& what) { fail(FS_E_RUNTIME, "Checkpoint: " + what); }

const fsjson::Value& member(const fsjson::Value& obj, const char* key)
{
    const fsjson::Value* v = obj.find(key);
    if (!v)
        bad_checkpoint(std::string("missing \"") + key + "\"");
    return *v;
}

double number_of(const fsjson::Value& v, const char* what)
{
    if (v.kind != fsjson::Value::Number || !std::isfinite(v.number))
        bad_checkpoint(std::string(what) + " must be a finite number");
    return v.number;
}

long long integer_of(const fsjson::Value& v, const char* what, long long lo, long long hi)
{
    const double x = number_of(v, what);
    if (x != std::floor(x) || x < (double)lo || x > (double)hi)
        bad_checkpoint(std::string(what) + " must be an integer in range");
    return (long long)x;
}

const std::string& string_of(const fsjson::Value& v, const char* what)
{
    if (v.kind != fsjson::Value::String)
        bad_checkpoint(std::string(what) + " must be a string");
    return v.text;
}

// Everything a load restores, fully validated before any live state changes.
struct CheckpointConfig {
    double ox = 0, oy = 0, gravity = 0, sigma = 0;
    fs_preconditioner precond = FS_PRECOND_MULTIGRID;
    fs_transport transport = FS_TRANSPORT_REFERENCE;
    std::vector<double> rho_l, rho_g, mu_l, mu_g;
    int predicted = 1;
};

CheckpointConfig parse_checkpoint(const fs_grid& g, const char* text, size_t length)
{
    const fsjson::Value meta = fsjson::parse(text, length);
    if (meta.kind != fsjson::Value::Object)
        bad_checkpoint("metadata must be a JSON object");
    const long long version = integer_of(member(meta, "version"), "version", -(1LL << 40), 1LL << 40);
    if (string_of(member(meta, "backend"), "backend") != kCheckpointBackend || version != kCheckpointVersion ||
        string_of(member(meta, "scalar"), "scalar") != "float64" ||
        string_of(member(meta, "byte_order"), "byte_order") != native_byte_order())
        bad_checkpoint("incompatible uniform HIP checkpoint format");
    const fsjson::Value& fields = member(meta, "payload");
    bool same_fields = fields.kind == fsjson::Value::Array && fields.items.size() == kCheckpointFieldCount;
    for (int i = 0; same_fields && i < kCheckpointFieldCount; ++i)
        same_fields = fields.items[i].kind == fsjson::Value::String && fields.items[i].text == kCheckpointFields[i];
    if (!same_fields)
        bad_checkpoint("incompatible payload layout");
    if (integer_of(member(meta, "width"), "width", 1, 1 << 30) != g.W ||
        integer_of(member(meta, "height"), "height", 1, 1 << 30) != g.H ||
        number_of(member(meta, "cell_size"), "cell_size") != g.h)
        bad_checkpoint("geometry does not match the fluid grid");

    CheckpointConfig c;
    const fsjson::Value& origin = member(meta, "origin");
    if (origin.kind != fsjson::Value::Array || origin.items.size() != 2)
        bad_checkpoint("origin must be a pair of numbers");
    c.ox = number_of(origin.items[0], "origin");
    c.oy = number_of(origin.items[1], "origin");
    c.gravity = number_of(member(meta, "gravity_y"), "gravity_y");
    c.sigma = number_of(member(meta, "surface_tension"), "surface_tension");
    if (c.sigma < 0.0)
        bad_checkpoint("invalid saved surface tension");
    const long long pc = integer_of(member(meta, "preconditioner"), "preconditioner", -1000, 1000);
    if (pc != FS_PRECOND_DIAGONAL && pc != FS_PRECOND_MULTIGRID)
        bad_checkpoint("unknown pressure preconditioner");
    c.precond = (fs_preconditioner)pc;
    const long long tr = integer_of(member(meta, "transport"), "transport", -1000, 1000);
    if (tr != FS_TRANSPORT_REFERENCE && tr != FS_TRANSPORT_NONE && tr != FS_TRANSPORT_SEMI_LAGRANGIAN)
        bad_checkpoint("unknown transport mode");
    c.transport = (fs_transport)tr;
    auto mixture = [&](const char* key, std::vector<double>& rho, std::vector<double>& mu) {
        const fsjson::Value& list = member(meta, key);
        if (list.kind != fsjson::Value::Array || list.items.size() > (size_t)kMaxComponents)
            bad_checkpoint(std::string(key) + " must be a list of at most 64 components");
        for (const fsjson::Value& e : list.items) {
            if (e.kind != fsjson::Value::Object)
                bad_checkpoint(std::string(key) + " entries must be objects");
            const double d = number_of(member(e, "density"), "density");
            if (!(d > 0.0))
                bad_checkpoint("component density must be positive");
            const fsjson::Value* m = e.find("dynamic_viscosity");
            const double v = m ? number_of(*m, "dynamic_viscosity") : 0.0;
            if (v < 0.0)
                bad_checkpoint("invalid saved material viscosity");
            rho.push_back(d);
            mu.push_back(v);
        }
    };
    mixture("liquid_mixture", c.rho_l, c.mu_l);
    mixture("gas_mixture", c.rho_g, c.mu_g);
    c.predicted = (int)integer_of(member(meta, "previous_solve_iterations"), "previous_solve_iterations", 0, 1 << 30);
    return c;
}

}  // namespace

extern "C" {

const char* fs_version(void) { return "fluidsim_hip 0.1 (gfx950)"; }

const char* fs_last_error(void) { return g_last_error.c_str(); }

fs_status fs_context_create(int device, fs_context** out) { return fs_context_create_ex(device, 0, out); }

fs_status fs_context_create_ex(int device, int flags, fs_context** out)
{
    return guard([&] {
        need((flags & ~(FS_CONTEXT_SINGLE_STREAM | FS_CONTEXT_PRIVATE_QUEUE)) == 0, FS_E_INVALID_ARGUMENT,
            "fs_context_create_ex: unknown flags");
        need(out != nullptr, FS_E_INVALID_ARGUMENT, "fs_context_create: null output pointer");
        int count = 0;
        HIPC(hipGetDeviceCount(&count));
        need(device >= 0 && device < count, FS_E_INVALID_ARGUMENT, "fs_context_create: no such HIP device");
        auto ctx = std::make_unique<fs_context>();
        ctx->device = device;
        HIPC(hipSetDevice(device));
        std::lock_guard<std::mutex> lock(device_setup_mutex());
        int lds = 0;
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
            ctx->lds_per_block = (size_t)lds;
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
            ctx->compute_units = cus;
        if ((flags & FS_CONTEXT_PRIVATE_QUEUE) && ctx->compute_units > 0) {
            // a CU mask over every compute unit: same placement as an unmasked stream, on a queue of its own
            const int cus = ctx->compute_units;
            const int words = (cus + 31) / 32;
            std::vector<uint32_t> mask((size_t)words, 0xffffffffu);
            if (cus % 32 != 0)
                mask.back() = (1u << (cus % 32)) - 1u;
            HIPC(hipExtStreamCreateWithCUMask(&ctx->stream, (uint32_t)words, mask.data()));
        } else {
            HIPC(hipStreamCreate(&ctx->stream));
        }
        if (!(flags & FS_CONTEXT_SINGLE_STREAM)) {
            HIPC(hipStreamCreate(&ctx->aux));
            HIPC(hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming));
            HIPC(hipEventCreateWithFlags(&ctx->join, hipEventDisableTiming));
            HIPC(hipEventCreateWithFlags(&ctx->join_levels, hipEventDisableTiming));
        }
        HIPC(hipHostMalloc((void**)&ctx->pinned, sizeof(PcgScalars), hipHostMallocDefault));
        HIPC(hipHostMalloc((void**)&ctx->pinned_value, sizeof(double), hipHostMallocDefault));
        viscosity_scratch_setup(device);   // this device's copy of the allocation-free solve's scratch
        HIPC(hipEventCreate(&ctx->ev0));
        HIPC(hipEventCreate(&ctx->ev1));
        *out = ctx.release();
    });
}

fs_status fs_context_destroy(fs_context* ctx)
{
    return guard([&] {
        if (!ctx)
            return;
        (void)hipSetDevice(ctx->device);
        if (ctx->stream)
            (void)hipStreamSynchronize(ctx->stream);
        {
            std::lock_guard<std::mutex> lock(device_setup_mutex());
            if (ctx->ev0)
                (void)hipEventDestroy(ctx->ev0);
            if (ctx->ev1)
                (void)hipEventDestroy(ctx->ev1);
            if (ctx->pinned)
                (void)hipHostFree(ctx->pinned);
            if (ctx->pinned_value)
                (void)hipHostFree(ctx->pinned_value);
            if (ctx->batch_pinned)
                (void)hipHostFree(ctx->batch_pinned);
            if (ctx->staging)
                (void)hipHostFree(ctx->staging);
            if (ctx->stream)
                (void)hipStreamDestroy(ctx->stream);
            if (ctx->aux)
                (void)hipStreamDestroy(ctx->aux);
            if (ctx->batch_stream) {
                (void)hipStreamSynchronize(ctx->batch_stream);
                (void)hipStreamDestroy(ctx->batch_stream);
            }
            if (ctx->batch_in)
                (void)hipEventDestroy(ctx->batch_in);
            if (ctx->batch_out)
                (void)hipEventDestroy(ctx->batch_out);
            if (ctx->fork)
                (void)hipEventDestroy(ctx->fork);
            if (ctx->join)
                (void)hipEventDestroy(ctx->join);
            if (ctx->join_levels)
                (void)hipEventDestroy(ctx->join_levels);
        }
        // outside the lock: the descriptor buffer's release takes it
        delete ctx;
    });
}

fs_status fs_context_synchronize(fs_context* ctx)
{
    return guard([&] {
        need(ctx != nullptr, FS_E_INVALID_ARGUMENT, "null context");
        set_device(ctx);
        HIPC(hipStreamSynchronize(ctx->stream));
    });
}

fs_status fs_grid_create(fs_context* ctx, int width, int height, double cell_size, fs_grid** out)
{
    return guard([&] {
        need(ctx != nullptr && out != nullptr, FS_E_INVALID_ARGUMENT, "fs_grid_create: null pointer");
        need(width > 0 && height > 0, FS_E_INVALID_ARGUMENT, "fs_grid_create: grid dimensions must be positive");
        need((long long)width * height <= (1LL << 28), FS_E_INVALID_ARGUMENT, "fs_grid_create: grid too large");
        need(std::isfinite(cell_size) && cell_size > 0.0, FS_E_INVALID_ARGUMENT,
            "fs_grid_create: cell size must be positive");
        set_device(ctx);
        auto g = std::make_unique<fs_grid>();
        g->ctx = ctx;
        g->W = width;
        g->H = height;
        g->h = cell_size;
        g->allocate();
        *out = g.release();
    });
}

fs_status fs_grid_destroy(fs_grid* grid)
{
    return guard([&] {
        if (!grid)
            return;
        set_device(grid->ctx);
        grid->drain();
        delete grid;
    });
}

fs_status fs_grid_shape(const fs_grid* grid, int* width, int* height, double* cell_size)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        if (width) *width = grid->W;
        if (height) *height = grid->H;
        if (cell_size) *cell_size = grid->h;
    });
}

fs_status fs_set_origin(fs_grid* grid, double x, double y)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        grid->ox = x;
        grid->oy = y;
    });
}

fs_status fs_get_origin(const fs_grid* grid, double* x, double* y)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        if (x) *x = grid->ox;
        if (y) *y = grid->oy;
    });
}

fs_status fs_configure_materials(fs_grid* grid, int n_liquid, const double* liquid_density,
    const double* liquid_viscosity, int n_gas, const double* gas_density, const double* gas_viscosity)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(n_liquid >= 0 && n_gas >= 0, FS_E_INVALID_ARGUMENT, "component counts must be non-negative");
        need((n_liquid == 0 || liquid_density) && (n_gas == 0 || gas_density), FS_E_INVALID_ARGUMENT,
            "component densities are required");
        need(n_liquid <= kMaxComponents && n_gas <= kMaxComponents, FS_E_INVALID_ARGUMENT,
            "at most 64 components per phase");
        for (int k = 0; k < n_liquid && liquid_viscosity; ++k)
            need(std::isfinite(liquid_viscosity[k]) && liquid_viscosity[k] >= 0.0, FS_E_INVALID_ARGUMENT,
                "Material viscosity must be finite and nonnegative");
        for (int k = 0; k < n_gas && gas_viscosity; ++k)
            need(std::isfinite(gas_viscosity[k]) && gas_viscosity[k] >= 0.0, FS_E_INVALID_ARGUMENT,
                "Material viscosity must be finite and nonnegative");
        grid->rho_l.assign(liquid_density, liquid_density + n_liquid);
        grid->rho_g.assign(gas_density, gas_density + n_gas);
        grid->mu_l.assign(n_liquid, 0.0);
        grid->mu_g.assign(n_gas, 0.0);
        for (int k = 0; k < n_liquid && liquid_viscosity; ++k)
            grid->mu_l[k] = liquid_viscosity[k];
        for (int k = 0; k < n_gas && gas_viscosity; ++k)
            grid->mu_g[k] = gas_viscosity[k];
        // Changing mixtures requires another upload (FluidGridBase.h:41).
        grid->uploaded = false;
        grid->has_snapshot = false;
    });
}

fs_status fs_set_gravity(fs_grid* grid, double gravity_y)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        grid->gravity_y = gravity_y;
    });
}

fs_status fs_get_gravity(const fs_grid* grid, double* gravity_y)
{
    return guard([&] {
        need(grid != nullptr && gravity_y != nullptr, FS_E_INVALID_ARGUMENT, "null pointer");
        *gravity_y = grid->gravity_y;
    });
}

fs_status fs_set_surface_tension(fs_grid* grid, double coefficient)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(std::isfinite(coefficient) && coefficient >= 0.0, FS_E_INVALID_ARGUMENT,
            "Surface tension must be finite and nonnegative");
        grid->sigma = coefficient;
    });
}

fs_status fs_set_solid_mask(fs_grid* grid, const unsigned char* mask)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        const int W = grid->W, H = grid->H;
        const size_t N = grid->N();
        std::vector<unsigned char> m;
        long long solids = 0;
        if (mask) {
            const std::string problem = solid_mask_problem(mask, W, H);
            need(problem.empty(), FS_E_INVALID_ARGUMENT, problem.c_str());
            m.resize(N);
            for (size_t i = 0; i < N; ++i) {
                m[i] = mask[i] ? 1 : 0;
                solids += m[i];
            }
        }
        set_device(grid->ctx);
        grid->drain();
        if (mask) {
            if (grid->solid.n != N)
                grid->bytes += grid->solid.alloc(N);
            device_upload(grid->solid.p, m.data(), N);
        } else if (grid->solid.n) {
            grid->bytes -= grid->solid.n;
            grid->solid.release();
        }
        grid->solid_cells = solids;
        grid->coeff_valid = false;
        grid->mg_valid = false;
        grid->drop_graph();
    });
}

fs_status fs_set_multigrid_precision(fs_grid* grid, fs_mg_precision mode)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(mode == FS_MG_FP32 || mode == FS_MG_FP16 || mode == FS_MG_BF16 || mode == FS_MG_BF16_VECTORS,
            FS_E_INVALID_ARGUMENT, "unknown multigrid precision");
        if (mode == grid->mg_prec)
            return;
        need(mode == FS_MG_FP32 || grid->smoother == FS_SMOOTHER_RED_BLACK, FS_E_UNSUPPORTED,
            "16-bit multigrid levels run the red-black smoother: select FS_SMOOTHER_RED_BLACK first");
        set_device(grid->ctx);
        grid->drain();
        grid->mg_prec = mode;
        grid->apply_precision();
        grid->tune_for_precision();
    });
}

fs_status fs_set_pressure_smoother(fs_grid* grid, fs_pressure_smoother kind, int sweeps, int coarse_sweeps)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(kind == FS_SMOOTHER_RED_BLACK || kind == FS_SMOOTHER_CHEBYSHEV, FS_E_INVALID_ARGUMENT,
            "unknown pressure smoother");
        need(sweeps >= 1 && sweeps <= (kind == FS_SMOOTHER_RED_BLACK ? 4 : kJacobiMaxSweeps), FS_E_INVALID_ARGUMENT,
            "sweeps must be 1..4 (red-black) or 1..8 (Chebyshev)");
        need(coarse_sweeps >= 1 && coarse_sweeps <= 4, FS_E_INVALID_ARGUMENT, "coarse sweeps must be 1..4");
        need(kind == FS_SMOOTHER_RED_BLACK || grid->mg_prec == FS_MG_FP32, FS_E_UNSUPPORTED,
            "the Chebyshev smoother runs FP32 multigrid levels: set the multigrid precision to FS_MG_FP32");
        if (kind == grid->smoother && sweeps == grid->smooth_sweeps && coarse_sweeps == grid->coarse_sweeps)
            return;
        set_device(grid->ctx);
        grid->drain();
        grid->set_smoother(kind, sweeps, coarse_sweeps);
    });
}

fs_status fs_set_viscosity_smoother(fs_grid* grid, int chebyshev_sweeps)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(chebyshev_sweeps == 0 || chebyshev_sweeps == 2 || chebyshev_sweeps == 4 || chebyshev_sweeps == 6,
            FS_E_INVALID_ARGUMENT, "viscosity smoother sweeps must be 0, 2, 4 or 6");
        if (chebyshev_sweeps == grid->vis_sweeps)
            return;
        set_device(grid->ctx);
        grid->drain();
        grid->vis_sweeps = chebyshev_sweeps;
        if (grid->visc)
            viscosity_set_smoother(grid->visc, chebyshev_sweeps);
    });
}

// The state a projection changes (velocities, both pressure histories, the iteration prediction), kept on
// the device so that autotuning trials start from the same state and leave it unchanged.
struct ProjectionState {
    DevBuf<double> u, v, p, pp;
    int predicted = 0;
    static void copy(hipStream_t st, double* dst, const double* src, size_t n)
    {
        if (n)
            HIPC(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    }
    void save(fs_grid& g)
    {
        u.alloc(g.u.n);
        v.alloc(g.v.n);
        p.alloc(g.p.n);
        pp.alloc(g.p_prev.n);
        copy(g.st(), u.p, g.u.p, g.u.n);
        copy(g.st(), v.p, g.v.p, g.v.n);
        copy(g.st(), p.p, g.p.p, g.p.n);
        copy(g.st(), pp.p, g.p_prev.p, g.p_prev.n);
        predicted = g.predicted;
    }
    void restore(fs_grid& g) const
    {
        copy(g.st(), g.u.p, u.p, g.u.n);
        copy(g.st(), g.v.p, v.p, g.v.n);
        copy(g.st(), g.p.p, p.p, g.p.n);
        copy(g.st(), g.p_prev.p, pp.p, g.p_prev.n);
        g.predicted = predicted;
    }
};

struct SmootherChoice {
    fs_pressure_smoother kind;
    int sweeps, coarse;
};

fs_status fs_tune_pressure_smoother(fs_grid* grid, double dt, int max_iterations, double divergence_tolerance,
    fs_pressure_smoother* kind, int* sweeps, int* coarse_sweeps, double* ms)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(grid->uploaded, FS_E_LOGIC, "Upload the grid state before tuning");
        need(std::isfinite(dt) && dt > 0.0, FS_E_INVALID_ARGUMENT, "dt must be finite and positive");
        fs_grid& g = *grid;
        set_device(g.ctx);
        g.drain();
        ProjectionState saved;
        saved.save(g);
        auto choose = [&](const SmootherChoice& c) {
            HIPC(hipStreamSynchronize(g.st()));
            g.set_smoother(c.kind, c.sweeps, c.coarse);
        };
        std::vector<SmootherChoice> candidates;
        for (int coarse : {4, 2}) {
            for (int n : {4, 3, 2})
                candidates.push_back({FS_SMOOTHER_RED_BLACK, n, coarse});
            if (g.mg_prec == FS_MG_FP32)
                for (int n : {3, 4, 5, 6, 8})
                    candidates.push_back({FS_SMOOTHER_CHEBYSHEV, n, coarse});
        }
        const SmootherChoice original{g.smoother, g.smooth_sweeps, g.coarse_sweeps};
        SmootherChoice best = original;
        double best_ms = std::numeric_limits<double>::infinity();
        try {
            for (const SmootherChoice& c : candidates) {
                choose(c);
                double fastest = std::numeric_limits<double>::infinity();
                bool converged = true;
                for (int rep = 0; rep < 3; ++rep) {   // the first is a warm-up (graph capture, first replays)
                    saved.restore(g);
                    HIPC(hipStreamSynchronize(g.st()));
                    const auto t0 = std::chrono::steady_clock::now();
                    const fs_projection_result res = g.project(dt, max_iterations, divergence_tolerance);
                    const double t = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                    converged = converged && res.converged;
                    if (rep > 0)
                        fastest = std::min(fastest, t);
                }
                if (converged && fastest < best_ms) {
                    best_ms = fastest;
                    best = c;
                }
            }
        } catch (...) {
            choose(original);
            saved.restore(g);
            HIPC(hipStreamSynchronize(g.st()));
            throw;
        }
        choose(best);
        saved.restore(g);
        HIPC(hipStreamSynchronize(g.st()));
        if (kind)
            *kind = best.kind;
        if (sweeps)
            *sweeps = best.sweeps;
        if (coarse_sweeps)
            *coarse_sweeps = best.coarse;
        if (ms)
            *ms = best_ms;
    });
}

fs_status fs_get_pressure_smoother(const fs_grid* grid, fs_pressure_smoother* kind, int* sweeps, int* coarse_sweeps)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        if (kind)
            *kind = grid->smoother;
        if (sweeps)
            *sweeps = grid->smooth_sweeps;
        if (coarse_sweeps)
            *coarse_sweeps = grid->coarse_sweeps;
    });
}

fs_status fs_set_flexible_cg(fs_grid* grid, int enabled)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        if (grid->flexible_cg != (enabled != 0)) {
            grid->flexible_cg = enabled != 0;
            grid->drop_graph();
        }
    });
}

fs_status fs_set_relative_tolerance(fs_grid* grid, int enabled)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        grid->relative_tol = enabled != 0;
    });
}

fs_status fs_get_relative_tolerance(const fs_grid* grid, int* enabled)
{
    return guard([&] {
        need(grid != nullptr && enabled != nullptr, FS_E_INVALID_ARGUMENT, "null pointer");
        *enabled = grid->relative_tol ? 1 : 0;
    });
}

fs_status fs_get_multigrid_precision(const fs_grid* grid, fs_mg_precision* mode)
{
    return guard([&] {
        need(grid != nullptr && mode != nullptr, FS_E_INVALID_ARGUMENT, "null pointer");
        *mode = grid->mg_prec;
    });
}

fs_status fs_solid_cells(const fs_grid* grid, long long* count)
{
    return guard([&] {
        need(grid != nullptr && count != nullptr, FS_E_INVALID_ARGUMENT, "null pointer");
        *count = grid->solid_cells;
    });
}

fs_status fs_capillary_step_limit(const fs_grid* grid, double* limit)
{
    return guard([&] {
        need(grid != nullptr && limit != nullptr, FS_E_INVALID_ARGUMENT, "null pointer");
        *limit = grid->capillary_step_limit();
    });
}

fs_status fs_set_preconditioner(fs_grid* grid, fs_preconditioner kind)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(kind == FS_PRECOND_DIAGONAL || kind == FS_PRECOND_MULTIGRID, FS_E_INVALID_ARGUMENT,
            "unknown preconditioner");
        grid->precond = kind;
    });
}

fs_status fs_set_transport(fs_grid* grid, fs_transport mode)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(mode == FS_TRANSPORT_REFERENCE || mode == FS_TRANSPORT_NONE || mode == FS_TRANSPORT_SEMI_LAGRANGIAN,
            FS_E_INVALID_ARGUMENT, "unknown transport mode");
        grid->transport = mode;
    });
}

fs_status fs_upload(fs_grid* grid, const double* liquid_fraction, const double* u, const double* v,
    const double* pressure, const double* liquid_mass, const double* gas_mass)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        set_device(grid->ctx);
        fs_grid& g = *grid;
        const size_t N = g.N();
        const size_t nl = g.rho_l.size(), ng = g.rho_g.size();
        g.drain();
        if (g.lmass.n != nl * N)
            g.bytes += g.lmass.alloc(nl * N);
        if (g.gmass.n != ng * N)
            g.bytes += g.gmass.alloc(ng * N);
        auto put = [&](DevBuf<double>& dst, const double* src, size_t n) {
            if (n == 0)
                return;
            if (src)
                HIPC(hipMemcpyAsync(dst.p, src, n * sizeof(double), hipMemcpyHostToDevice, g.st()));
            else
                HIPC(hipMemsetAsync(dst.p, 0, n * sizeof(double), g.st()));
        };
        put(g.fraction, liquid_fraction, N);
        put(g.u, u, g.NU());
        put(g.v, v, g.NV());
        put(g.lmass, liquid_mass, nl * N);
        put(g.gmass, gas_mass, ng * N);
        if (pressure) {
            HIPC(hipMemcpyAsync(g.tmp64.p, pressure, N * sizeof(double), hipMemcpyHostToDevice, g.st()));
            launch_pad_copy(g.st(), g.tmp64.p, g.p.p, g.W, g.H, 1);
        } else {
            HIPC(hipMemsetAsync(g.p.p, 0, g.NP() * sizeof(double), g.st()));
        }
        // upload resets the pressure history (FluidGridCPU.cpp:249)
        HIPC(hipMemcpyAsync(g.p_prev.p, g.p.p, g.NP() * sizeof(double), hipMemcpyDeviceToDevice, g.st()));
        HIPC(hipStreamSynchronize(g.st()));
        g.uploaded = true;
        g.coeff_valid = false;
        g.mg_valid = false;
    });
}

fs_status fs_download(fs_grid* grid, double* liquid_fraction, double* u, double* v, double* pressure,
    double* liquid_mass, double* gas_mass)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(grid->uploaded, FS_E_LOGIC, "Upload the grid state before downloading it");
        set_device(grid->ctx);
        fs_grid& g = *grid;
        const size_t N = g.N();
        auto get = [&](double* dst, const DevBuf<double>& src, size_t n) {
            if (dst && n)
                HIPC(hipMemcpyAsync(dst, src.p, n * sizeof(double), hipMemcpyDeviceToHost, g.st()));
        };
        get(liquid_fraction, g.fraction, N);
        get(u, g.u, g.NU());
        get(v, g.v, g.NV());
        get(liquid_mass, g.lmass, g.rho_l.size() * N);
        get(gas_mass, g.gmass, g.rho_g.size() * N);
        if (pressure) {
            launch_pad_copy(g.st(), g.p.p, g.tmp64.p, g.W, g.H, 0);
            HIPC(hipMemcpyAsync(pressure, g.tmp64.p, N * sizeof(double), hipMemcpyDeviceToHost, g.st()));
        }
        HIPC(hipStreamSynchronize(g.st()));
    });
}

fs_status fs_invalidate_caches(fs_grid* grid)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        grid->coeff_valid = false;
        grid->mg_valid = false;
    });
}

// writeColorPixels (FluidGridBase.h:139, FluidGridCPU.cpp:1826-1860)
fs_status fs_color(fs_grid* grid, const float* component_rgb, int liquid_fraction_alpha, float* pixels)
{
    return guard([&] {
        need(grid != nullptr && pixels != nullptr, FS_E_INVALID_ARGUMENT, "null pointer");
        fs_grid& g = *grid;
        need(g.uploaded, FS_E_LOGIC, "Upload the grid state before rendering it");
        const size_t nl = g.rho_l.size(), ng = g.rho_g.size(), N = g.N();
        need(g.lmass.n == nl * N && g.gmass.n == ng * N, FS_E_LOGIC,
            "Upload the configured fluid mixtures before rendering them");
        need(component_rgb != nullptr || nl + ng == 0, FS_E_INVALID_ARGUMENT, "null component colours");
        ColorTable table{};
        for (size_t k = 0; k < nl + ng; ++k) {
            for (int c = 0; c < 3; ++c)
                table.rgb[k][c] = component_rgb[3 * k + c];
            table.density[k] = k < nl ? g.rho_l[k] : g.rho_g[k - nl];
        }
        set_device(g.ctx);
        if (g.color_px.n != 4 * N)
            g.bytes += g.color_px.alloc(4 * N);
        launch_color(g.st(), g.dev(), table, liquid_fraction_alpha != 0, g.color_px.p);
        HIPC(hipMemcpyAsync(pixels, g.color_px.p, 4 * N * sizeof(float), hipMemcpyDeviceToHost, g.st()));
        HIPC(hipStreamSynchronize(g.st()));
    });
}

// commitEdits (FluidGridBase.h:142, FluidGridCPU.cpp:253-284) on resolved edits; duplicates are reduced as
// CellEditBatch::prepare does (FluidGridEdit.h:37-95): raster order, the last edit of a cell wins.
fs_status fs_commit_edits(fs_grid* grid, int count, const fs_cell_edit* cells, const double* velocity)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(velocity == nullptr, FS_E_UNSUPPORTED, "Material edit velocity requires uniform WebGPU");
        need(count >= 0 && (count == 0 || cells != nullptr), FS_E_INVALID_ARGUMENT, "null edit list");
        fs_grid& g = *grid;
        const int components = (int)(g.rho_l.size() + g.rho_g.size());
        std::vector<EditCell> list((size_t)count);
        for (int i = 0; i < count; ++i) {
            const fs_cell_edit& c = cells[i];
            need(c.x >= 0 && c.x < g.W && c.y >= 0 && c.y < g.H, FS_E_INVALID_ARGUMENT,
                "Fluid grid edit cell is outside the grid");
            need(c.component >= 0 && c.component < components, FS_E_INVALID_ARGUMENT,
                "Fluid grid edit component is not configured");
            need(std::isfinite(c.mass) && c.mass >= 0.0, FS_E_INVALID_ARGUMENT,
                "Fluid grid edit mass must be finite and non-negative");
            list[(size_t)i] = EditCell{c.x, c.y, c.component, 0, c.mass};
        }
        const size_t N = g.N();
        need(g.uploaded && g.lmass.n == g.rho_l.size() * N && g.gmass.n == g.rho_g.size() * N, FS_E_LOGIC,
            "Upload the configured fluid mixtures before committing edits");
        if (count == 0)
            return;
        std::stable_sort(list.begin(), list.end(), [](const EditCell& a, const EditCell& b) {
            return a.y < b.y || (a.y == b.y && a.x < b.x);
        });
        size_t kept = 0;
        for (size_t first = 0; first < list.size();) {
            size_t last = first + 1;
            while (last < list.size() && list[last].x == list[first].x && list[last].y == list[first].y)
                ++last;
            list[kept++] = list[last - 1];
            first = last;
        }
        list.resize(kept);
        set_device(g.ctx);
        g.drain();
        if (g.edit_cells.n < kept)
            g.bytes += g.edit_cells.alloc(kept);
        const SyncOnExit drain{g.st()};
        stage_upload(g.ctx, g.edit_cells.p, list.data(), kept * sizeof(EditCell));
        launch_commit_edits(g.st(), g.dev(), g.edit_cells.p, (int)kept);
        HIPC(hipStreamSynchronize(g.st()));
        g.coeff_valid = false;   // invalidateCaches (FluidGridCPU.cpp:282)
        g.mg_valid = false;
    });
}

fs_status fs_project(fs_grid* grid, double dt, int max_iterations, double divergence_tolerance,
    fs_projection_result* result)
{
    return guard([&] {
        need(grid != nullptr && result != nullptr, FS_E_INVALID_ARGUMENT, "null pointer");
        set_device(grid->ctx);
        *result = grid->project(dt, max_iterations, divergence_tolerance);
    });
}

fs_status fs_apply_gravity(fs_grid* grid, double dt)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(grid->uploaded, FS_E_LOGIC, "Upload the grid state before applying forces");
        set_device(grid->ctx);
        launch_gravity(grid->st(), grid->dev(), grid->gravity_y, dt);
        HIPC(hipStreamSynchronize(grid->st()));
    });
}

fs_status fs_add_vortex(fs_grid* grid, double cx, double cy, double radius, double strength)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(grid->uploaded, FS_E_LOGIC, "Upload the grid state before editing velocities");
        set_device(grid->ctx);
        launch_add_vortex(grid->st(), grid->dev(), grid->ox, grid->oy, cx, cy, radius, strength);
        HIPC(hipStreamSynchronize(grid->st()));
    });
}

fs_status fs_max_divergence(fs_grid* grid, double* out)
{
    return guard([&] {
        need(grid != nullptr && out != nullptr, FS_E_INVALID_ARGUMENT, "null pointer");
        need(grid->uploaded, FS_E_LOGIC, "Upload the grid state before measuring divergence");
        set_device(grid->ctx);
        *out = grid->max_divergence();
    });
}

static void advect_sl(fs_grid& g, double dt)
{
    const size_t N = g.N();
    const size_t fields = 1 + g.rho_l.size() + g.rho_g.size();
    if (g.sl_u.n != g.NU())
        g.bytes += g.sl_u.alloc(g.NU());
    if (g.sl_v.n != g.NV())
        g.bytes += g.sl_v.alloc(g.NV());
    if (g.sl_cells.n != fields * N)
        g.bytes += g.sl_cells.alloc(fields * N);
    launch_semi_lagrangian(g.st(), g.dev(), g.ox, g.oy, dt, g.sl_u.p, g.sl_v.p, g.sl_cells.p);
    HIPC(hipMemcpyAsync(g.u.p, g.sl_u.p, g.NU() * sizeof(double), hipMemcpyDeviceToDevice, g.st()));
    HIPC(hipMemcpyAsync(g.v.p, g.sl_v.p, g.NV() * sizeof(double), hipMemcpyDeviceToDevice, g.st()));
    HIPC(hipMemcpyAsync(g.fraction.p, g.sl_cells.p, N * sizeof(double), hipMemcpyDeviceToDevice, g.st()));
    if (g.rho_l.size())
        HIPC(hipMemcpyAsync(g.lmass.p, g.sl_cells.p + N, g.rho_l.size() * N * sizeof(double),
            hipMemcpyDeviceToDevice, g.st()));
    if (g.rho_g.size())
        HIPC(hipMemcpyAsync(g.gmass.p, g.sl_cells.p + (1 + g.rho_l.size()) * N, g.rho_g.size() * N * sizeof(double),
            hipMemcpyDeviceToDevice, g.st()));
    g.coeff_valid = false;
    g.mg_valid = false;
}

// Most CFL substeps one transport call may take; beyond it the velocities are unphysical (the reference
// would loop for as long as the ratio says, or forever at an infinite speed).
constexpr double kMaxTransportSubsteps = 1e6;

// FluidGridCPU::advect (FluidGridCPU.cpp:362-463).  Returns the number of CFL substeps.
static int advect_reference(fs_grid& g, double dt)
{
    const size_t N = g.N(), NU = g.NU(), NV = g.NV();
    const size_t nl = g.rho_l.size(), ng = g.rho_g.size();
    auto ensure = [&](auto& b, size_t n) {
        if (b.n != n || (n && !b.p))
            g.bytes += b.alloc(n);
    };
    ensure(g.tr_fraction, N);
    ensure(g.tr_lmass, nl * N);
    ensure(g.tr_gmass, ng * N);
    for (int k = 0; k < 2; ++k) {
        ensure(g.tr_nu[k], NU);
        ensure(g.tr_nv[k], NV);
    }
    ensure(g.tr_mom_u, NU);
    ensure(g.tr_mom_v, NV);
    ensure(g.tr_lt, std::max(NU, NV));
    ensure(g.tr_mt, std::max(NU, NV));
    ensure(g.tr_lcw, nl * N);
    ensure(g.tr_gcw, ng * N);
    ensure(g.tr_mcw, N);
    ensure(g.tr_split, N);
    ensure(g.tr_out, 1);

    int cur = 0;   // tr_nu[cur] / tr_nv[cur] hold the transported velocities
    auto view = [&]() {
        TransportDev t{};
        t.W = g.W;
        t.H = g.H;
        t.nl = (int)nl;
        t.ng = (int)ng;
        t.icv = 1.0 / (g.h * g.h);
        t.u = g.u.p;
        t.v = g.v.p;
        t.f = g.fraction.p;
        t.f_out = g.tr_fraction.p;
        t.lm = g.lmass.p;
        t.lm_out = g.tr_lmass.p;
        t.gm = g.gmass.p;
        t.gm_out = g.tr_gmass.p;
        t.density = g.density.p;
        t.nu_in = g.tr_nu[cur].p;
        t.nu_out = g.tr_nu[cur ^ 1].p;
        t.nv_in = g.tr_nv[cur].p;
        t.nv_out = g.tr_nv[cur ^ 1].p;
        t.mom_u = g.tr_mom_u.p;
        t.mom_v = g.tr_mom_v.p;
        t.lt = g.tr_lt.p;
        t.mt = g.tr_mt.p;
        t.lcw = g.tr_lcw.p;
        t.gcw = g.tr_gcw.p;
        t.mcw = g.tr_mcw.p;
        t.split = g.tr_split.p;
        t.partials = g.partials.p;
        t.out = g.tr_out.p;
        return t;
    };
    auto swap_state = [&]() {
        g.fraction.swap(g.tr_fraction);
        g.lmass.swap(g.tr_lmass);
        g.gmass.swap(g.tr_gmass);
        cur ^= 1;
    };

    // rebuildMomentum into tr_nu[0] / tr_nv[0], then the CFL limit
    launch_density(g.st(), g.dev());
    cur = 1;
    launch_transport_begin(g.st(), view());
    cur = 0;
    HIPC(hipMemcpyAsync(g.ctx->pinned_value, g.tr_out.p, sizeof(double), hipMemcpyDeviceToHost, g.st()));
    HIPC(hipStreamSynchronize(g.st()));
    const double max_outflow = *g.ctx->pinned_value;
    // no cell exports more than half its volume per substep (FluidGridCPU.cpp:388-389)
    const double max_step = max_outflow > 0.0 ? 0.5 * g.h / max_outflow : dt;
    if (!(max_step > 0.0) || dt / max_step > kMaxTransportSubsteps)
        fail(FS_E_RUNTIME, "Transport would need more than 1e6 CFL substeps (non-finite or extreme velocity)");

    double remaining = dt;
    int substeps = 0;
    while (remaining > 0.0) {
        const double step = std::min(remaining, max_step);
        const double scale = step / g.h;
        launch_transport_freeze(g.st(), view());
        launch_transport_pass(g.st(), view(), 1, scale);
        swap_state();
        launch_transport_pass(g.st(), view(), 0, scale);
        swap_state();
        remaining -= step;
        ++substeps;
    }
    g.u.swap(g.tr_nu[cur]);
    g.v.swap(g.tr_nv[cur]);
    g.coeff_valid = false;
    g.mg_valid = false;
    return substeps;
}

// Iteration budget standing in for the reference's unbounded viscosity solve (maximum = -1).
constexpr int kViscosityBudget = 100000;

static void raise_viscosity(int status, double relative)
{
    if (status == 1)
        fail(FS_E_RUNTIME, "Viscosity received nonfinite velocity/material state");
    if (status == 2)
        fail(FS_E_RUNTIME, "Implicit viscosity solver breakdown");
    if (status == 3)
        fail(FS_E_RUNTIME, "Implicit viscosity did not converge; relative residual " + std::to_string(relative));
}

// The volume-weighted (rho, mu) material of FluidGridGPU::applyViscosity (FluidGridGPU.cpp:995-1031),
// blended on the grid's stream into the solver's material buffer.
static double* blend_viscosity_material(fs_grid& g)
{
    if (!g.visc) {
        g.visc = viscosity_create(g.W, g.H);
        g.bytes += viscosity_bytes(g.visc);
        viscosity_set_smoother(g.visc, g.vis_sweeps);
    }
    VisMixture mix{};
    for (size_t k = 0; k < g.rho_l.size(); ++k) {
        mix.density[k] = g.rho_l[k];
        mix.viscosity[k] = g.mu_l[k];
    }
    for (size_t k = 0; k < g.rho_g.size(); ++k) {
        mix.density[kMaxComponents + k] = g.rho_g[k];
        mix.viscosity[kMaxComponents + k] = g.mu_g[k];
    }
    double* material = viscosity_material_buffer(g.visc);
    launch_viscosity_blend(g.st(), (int)g.N(), g.h, (int)g.rho_l.size(), (int)g.rho_g.size(), g.lmass.p,
        g.gmass.p, mix, material);
    return material;
}

// FluidGridGPU::applyViscosity (FluidGridGPU.cpp:987-1046): blend, then solve to 1e-4.  Returns the
// dispatched iteration count.
static int apply_viscosity(fs_grid& g, double dt)
{
    if (dt <= 0.0 || !g.has_viscosity())
        return 0;
    double* material = blend_viscosity_material(g);
    int iterations = 0;
    double relative = 0.0;
    const int status = viscosity_solve(g.visc, g.st(), g.u.p, g.v.p, material, dt, g.h, 1e-4, kViscosityBudget,
        &iterations, &relative);
    raise_viscosity(status, relative);
    return iterations;
}

fs_status fs_apply_viscosity(fs_grid* grid, double dt, int* iterations)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(grid->uploaded, FS_E_LOGIC, "Upload the grid state before applying viscosity");
        set_device(grid->ctx);
        const int n = apply_viscosity(*grid, dt);
        HIPC(hipStreamSynchronize(grid->st()));
        if (iterations)
            *iterations = n;
    });
}

fs_status fs_tune_viscosity_smoother(fs_grid* grid, double dt, int* chebyshev_sweeps, double* ms)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(grid->uploaded, FS_E_LOGIC, "Upload the grid state before tuning");
        need(std::isfinite(dt) && dt > 0.0, FS_E_INVALID_ARGUMENT, "dt must be finite and positive");
        fs_grid& g = *grid;
        need(g.has_viscosity(), FS_E_LOGIC, "The grid has no viscosity to tune");
        set_device(g.ctx);
        g.drain();
        ProjectionState saved;
        saved.save(g);
        auto choose = [&](int n) {
            HIPC(hipStreamSynchronize(g.st()));
            g.vis_sweeps = n;
            if (g.visc)
                viscosity_set_smoother(g.visc, n);
        };
        const int original = g.vis_sweeps;
        int best = original;
        double best_ms = std::numeric_limits<double>::infinity();
        try {
            for (int n : {0, 2, 4, 6}) {
                choose(n);
                double fastest = std::numeric_limits<double>::infinity();
                bool converged = true;
                for (int rep = 0; rep < 3; ++rep) {   // the first is a warm-up (graph capture, first replays)
                    saved.restore(g);
                    HIPC(hipStreamSynchronize(g.st()));
                    const auto t0 = std::chrono::steady_clock::now();
                    double* material = blend_viscosity_material(g);   // creates the solver with g.vis_sweeps
                    int iterations = 0;
                    double relative = 0.0;
                    const int status = viscosity_solve(g.visc, g.st(), g.u.p, g.v.p, material, dt, g.h, 1e-4,
                        kViscosityBudget, &iterations, &relative);
                    HIPC(hipStreamSynchronize(g.st()));
                    const double t = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                    converged = converged && status != 1 && status != 2 && status != 3;
                    if (rep > 0)
                        fastest = std::min(fastest, t);
                }
                if (converged && fastest < best_ms) {
                    best_ms = fastest;
                    best = n;
                }
            }
        } catch (...) {
            choose(original);
            saved.restore(g);
            HIPC(hipStreamSynchronize(g.st()));
            throw;
        }
        choose(best);
        saved.restore(g);
        HIPC(hipStreamSynchronize(g.st()));
        if (chebyshev_sweeps)
            *chebyshev_sweeps = best;
        if (ms)
            *ms = best_ms;
    });
}

// Every grid of a batch shares `ctx`, is uploaded and appears once.
static fs_context* batch_context(int count, fs_grid* const* grids, const char* upload_message)
{
    fs_context* ctx = grids[0] ? grids[0]->ctx : nullptr;
    for (int i = 0; i < count; ++i) {
        need(grids[i] != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(grids[i]->ctx == ctx, FS_E_INVALID_ARGUMENT, "batched grids must share one context");
        need(grids[i]->uploaded, FS_E_LOGIC, upload_message);
        for (int j = 0; j < i; ++j)
            need(grids[j] != grids[i], FS_E_INVALID_ARGUMENT, "a grid appears twice in the batch");
    }
    return ctx;
}

// The batched viscosity step of the grids listed in `which` (each with its own dt), one persistent launch
// for those at or under the context's split.  A grid's failure lands in errors[i] (its velocities unchanged)
// and does not stop the others.  Returns with the context's stream drained.
static void batch_viscosity(fs_context* ctx, fs_grid* const* grids, const std::vector<int>& which, const double* dts,
    int* iterations, std::vector<std::string>& errors)
{
    hipStream_t st = ctx->stream;
    const SyncOnExit drain{st};   // the staged descriptors and host results stay valid until then
    const size_t bytes = viscosity_batch_desc_bytes();
    std::vector<unsigned char> descs;
    std::vector<int> batched, alone, solo;
    for (int i : which) {
        fs_grid& g = *grids[i];
        const double dt = dts[i];
        if (dt <= 0.0 || !g.has_viscosity())
            continue;
        if (batch_solo(ctx, g.W, g.H)) {
            solo.push_back(i);
            continue;
        }
        double* material = blend_viscosity_material(g);
        descs.resize((batched.size() + 1) * bytes);
        if (viscosity_batch_prepare(g.visc, st, g.u.p, g.v.p, material, dt, g.h, 1e-4, kViscosityBudget,
                descs.data() + batched.size() * bytes))
            batched.push_back(i);
        else
            alone.push_back(i);
    }
    if (!batched.empty()) {
        const size_t need_bytes = batched.size() * bytes;
        if (ctx->vis_batch_desc.n < need_bytes)
            ctx->vis_batch_desc.alloc(need_bytes);
        stage_upload(ctx, ctx->vis_batch_desc.p, descs.data(), need_bytes);
        const bool spread = (int)batched.size() <= ctx->compute_units && ctx->lds_per_block >= kBatchSpreadLds;
        const hipStream_t bs = solo.empty() ? st : batch_side_stream(ctx);
        launch_batch_viscosity(bs, ctx->vis_batch_desc.p, (int)batched.size(), spread ? kBatchSpreadLds : 0);
        for (int i : batched)
            viscosity_batch_fetch(grids[i]->visc, bs);
        if (!solo.empty())
            HIPC(hipEventRecord(ctx->batch_out, bs));
    }
    auto attempt = [&](int i, auto&& body) {
        try {
            body();
        } catch (const std::exception& e) {
            errors[(size_t)i] = e.what();
        }
    };
    // grids above the batch split: their own solve, beside the persistent launch
    for (int i : solo)
        attempt(i, [&] {
            const int it = apply_viscosity(*grids[i], dts[i]);
            if (iterations)
                iterations[i] = it;
        });
    if (!solo.empty() && !batched.empty())
        HIPC(hipStreamWaitEvent(st, ctx->batch_out, 0));
    HIPC(hipStreamSynchronize(st));
    for (int i : batched) {
        int it = 0;
        double relative = 0.0;
        const int status = viscosity_batch_result(grids[i]->visc, kViscosityBudget, &it, &relative);
        if (status == 4) {
            alone.push_back(i);   // more iterations than the batch runs: solved again on its own
            continue;
        }
        if (iterations)
            iterations[i] = it;
        attempt(i, [&] { raise_viscosity(status, relative); });
    }
    for (int i : alone)
        attempt(i, [&] {
            const int it = apply_viscosity(*grids[i], dts[i]);
            if (iterations)
                iterations[i] = it;
        });
    HIPC(hipStreamSynchronize(st));
}

fs_status fs_batch_apply_viscosity(int count, fs_grid* const* grids, double dt, int* iterations)
{
    return guard([&] {
        need(count >= 0 && (count == 0 || grids), FS_E_INVALID_ARGUMENT, "null pointer");
        if (count == 0)
            return;
        fs_context* ctx = batch_context(count, grids, "Upload the grid state before applying viscosity");
        if (iterations)
            std::fill(iterations, iterations + count, 0);
        set_device(ctx);
        std::vector<int> all((size_t)count);
        for (int i = 0; i < count; ++i)
            all[(size_t)i] = i;
        const std::vector<double> dts((size_t)count, dt);
        std::vector<std::string> errors((size_t)count);
        batch_viscosity(ctx, grids, all, dts.data(), iterations, errors);
        for (const std::string& e : errors)
            if (!e.empty())
                fail(FS_E_RUNTIME, e);
    });
}

fs_status fs_set_component_viscosity(fs_grid* grid, int phase, int component, double coefficient)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(std::isfinite(coefficient) && coefficient >= 0.0, FS_E_INVALID_ARGUMENT,
            "Material viscosity must be finite and nonnegative");
        need(phase == 0 || phase == 1, FS_E_INVALID_ARGUMENT, "phase must be 0 (liquid) or 1 (gas)");
        std::vector<double>& mu = phase == 0 ? grid->mu_l : grid->mu_g;
        need(component >= 0 && (size_t)component < mu.size(), FS_E_INVALID_ARGUMENT, "no such mixture component");
        mu[(size_t)component] = coefficient;
    });
}

fs_status fs_viscosity_solve(fs_context* ctx, int width, int height, double cell_size, const double* material,
    double* u, double* v, double dt, double tolerance, int max_iterations, int* iterations,
    double* relative_residual)
{
    return guard([&] {
        need(ctx != nullptr && material != nullptr && u != nullptr && v != nullptr, FS_E_INVALID_ARGUMENT,
            "null pointer");
        need(width > 0 && height > 0 && cell_size > 0.0, FS_E_INVALID_ARGUMENT, "invalid grid shape");
        need(std::isfinite(tolerance) && tolerance >= 0.0, FS_E_INVALID_ARGUMENT, "invalid tolerance");
        set_device(ctx);
        const size_t N = (size_t)width * height, NU = (size_t)(width + 1) * height, NV = (size_t)width * (height + 1);
        std::unique_ptr<ViscositySolver, void (*)(ViscositySolver*)> solver(viscosity_create(width, height),
            viscosity_destroy);
        DevBuf<double> dm, du, dv;
        dm.alloc(2 * N);
        du.alloc(NU);
        dv.alloc(NV);
        HIPC(hipMemcpyAsync(dm.p, material, 2 * N * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        HIPC(hipMemcpyAsync(du.p, u, NU * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        HIPC(hipMemcpyAsync(dv.p, v, NV * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        int it = 0;
        double rel = 0.0;
        const int budget = max_iterations < 0 ? kViscosityBudget : max_iterations;
        const int status = viscosity_solve(solver.get(), ctx->stream, du.p, dv.p, dm.p, dt, cell_size, tolerance,
            budget, &it, &rel);
        if (iterations)
            *iterations = it;
        if (relative_residual)
            *relative_residual = rel;
        raise_viscosity(status, rel);
        HIPC(hipMemcpyAsync(u, du.p, NU * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIPC(hipMemcpyAsync(v, dv.p, NV * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIPC(hipStreamSynchronize(ctx->stream));
    });
}

fs_status fs_viscosity_solve_noalloc(fs_context* ctx, int width, int height, double cell_size,
    const double* material, double* u, double* v, double dt, double tolerance, int max_iterations, int* iterations,
    double* relative_residual)
{
    return guard([&] {
        need(ctx != nullptr && material != nullptr && u != nullptr && v != nullptr, FS_E_INVALID_ARGUMENT,
            "null pointer");
        need(width > 0 && height > 0 && cell_size > 0.0, FS_E_INVALID_ARGUMENT, "invalid grid shape");
        need(std::isfinite(tolerance) && tolerance >= 0.0, FS_E_INVALID_ARGUMENT, "invalid tolerance");
        need((long long)width * height <= (1LL << 26) && viscosity_scratch_bytes(width, height) <= kViscosityScratchBytes,
            FS_E_INVALID_ARGUMENT, "grid too large for the allocation-free viscosity solve's 64 MiB scratch");
        set_device(ctx);
        int it = 0;
        double rel = 0.0;
        const int budget = max_iterations < 0 ? kViscosityBudget : max_iterations;
        const int status = viscosity_solve_scratch(ctx->device, ctx->stream, width, height, u, v, material, dt,
            cell_size, tolerance, budget, &it, &rel);
        if (iterations)
            *iterations = it;
        if (relative_residual)
            *relative_residual = rel;
        raise_viscosity(status, rel);
    });
}

fs_status fs_advect(fs_grid* grid, double dt, int* substeps)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(grid->uploaded, FS_E_LOGIC, "Upload the grid state before transport");
        set_device(grid->ctx);
        const int n = dt > 0.0 ? advect_reference(*grid, dt) : 0;
        HIPC(hipStreamSynchronize(grid->st()));
        if (substeps)
            *substeps = n;
    });
}

fs_status fs_advect_semi_lagrangian(fs_grid* grid, double dt)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(grid->uploaded, FS_E_LOGIC, "Upload the grid state before transport");
        set_device(grid->ctx);
        advect_sl(*grid, dt);
        HIPC(hipStreamSynchronize(grid->st()));
    });
}

fs_status fs_step(fs_grid* grid, double dt, int max_iterations, double divergence_tolerance,
    fs_step_timings* timings, fs_projection_result* result)
{
    return guard([&] {
        need(grid != nullptr && result != nullptr, FS_E_INVALID_ARGUMENT, "null pointer");
        fs_grid& g = *grid;
        if (timings)
            *timings = fs_step_timings{};
        fs_projection_result res{};
        res.max_divergence = std::numeric_limits<double>::infinity();
        // FluidGridCPU::step returns {0, inf, false, 0, 0} for dt <= 0 (FluidGridCPU.cpp:306-307)
        if (dt <= 0.0) {
            res.physical_steps = 0;
            *result = res;
            return;
        }
        need(g.uploaded, FS_E_LOGIC, "Upload the grid state before stepping");
        set_device(g.ctx);
        using Clock = std::chrono::steady_clock;
        auto t0 = Clock::now();
        auto lap = [&](double* slot) {
            if (!slot)
                return;
            HIPC(hipStreamSynchronize(g.st()));
            const auto now = Clock::now();
            *slot += std::chrono::duration<double, std::milli>(now - t0).count();
            t0 = now;
        };
        // capillary subdivision (FluidGridCPU.cpp:319-357): count substeps of dt / count, each one
        // transport, viscosity (the WebGPU order, FluidGridGPU.cpp:1134-1158), gravity, surface
        // tension, projection; stop after the first substep that does not converge.
        const bool tension = g.sigma > 0.0;
        const double subdivisions = tension ? std::max(1.0, std::ceil(dt / g.capillary_step_limit())) : 1.0;
        if (subdivisions > (double)std::numeric_limits<int>::max())
            fail(FS_E_RUNTIME, "Surface tension requires too many physical substeps for this timestep");
        const int count = (int)subdivisions;
        const double step = dt / count;
        res.physical_steps = 0;
        for (int i = 0; i < count; ++i) {
            if (g.transport == FS_TRANSPORT_REFERENCE) {
                advect_reference(g, step);   // UniformCPU reports no transport subdivision count
            } else if (g.transport == FS_TRANSPORT_SEMI_LAGRANGIAN) {
                advect_sl(g, step);
                res.transport_steps += 1;
            }
            lap(timings ? &timings->transport_ms : nullptr);
            const int viscosity_iterations = apply_viscosity(g, step);
            res.viscosity_iterations += viscosity_iterations;
            if (viscosity_iterations)
                lap(timings ? &timings->viscosity_ms : nullptr);
            launch_gravity(g.st(), g.dev(), g.gravity_y, step);
            if (tension)
                g.apply_surface_tension(step);
            lap(timings ? &timings->forces_ms : nullptr);
            const fs_projection_result proj = g.project(step, max_iterations, divergence_tolerance);
            lap(timings ? &timings->pressure_ms : nullptr);
            res.iterations += proj.iterations;
            res.max_divergence = proj.max_divergence;
            res.converged = proj.converged;
            res.restarts = proj.restarts;
            res.initial_residual = proj.initial_residual;
            res.final_residual = proj.final_residual;
            res.rhs_max = proj.rhs_max;
            ++res.physical_steps;
            if (!res.converged)
                break;
        }
        *result = res;
    });
}

fs_status fs_apply_surface_tension(fs_grid* grid, double dt)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(grid->uploaded, FS_E_LOGIC, "Upload the grid state before applying surface tension");
        set_device(grid->ctx);
        if (dt > 0.0 && grid->sigma > 0.0)
            grid->apply_surface_tension(dt);
        HIPC(hipStreamSynchronize(grid->st()));
    });
}

fs_status fs_capillary_geometry(fs_grid* grid, double* heights, double* curvature)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(grid->uploaded, FS_E_LOGIC, "Upload the grid state first");
        set_device(grid->ctx);
        fs_grid& g = *grid;
        const size_t n = g.N();
        if (g.st_curv.n != n) {
            g.bytes += g.st_heights.alloc(n);
            g.bytes += g.st_orient.alloc(n);
            g.bytes += g.st_raw.alloc(n);
            g.bytes += g.st_curv.alloc(n);
        }
        launch_capillary_geometry(g.st(), g.W, g.H, g.h, g.fraction.p, g.st_heights.p, g.st_orient.p, g.st_raw.p,
            g.st_curv.p);
        std::vector<double2> hv(n), ov(n), cv(n);
        HIPC(hipMemcpyAsync(hv.data(), g.st_heights.p, n * sizeof(double2), hipMemcpyDeviceToHost, g.st()));
        HIPC(hipMemcpyAsync(ov.data(), g.st_orient.p, n * sizeof(double2), hipMemcpyDeviceToHost, g.st()));
        HIPC(hipMemcpyAsync(cv.data(), g.st_curv.p, n * sizeof(double2), hipMemcpyDeviceToHost, g.st()));
        HIPC(hipStreamSynchronize(g.st()));
        for (size_t i = 0; i < n; ++i) {
            if (heights) {
                heights[4 * i] = hv[i].x;
                heights[4 * i + 1] = hv[i].y;
                heights[4 * i + 2] = ov[i].x;
                heights[4 * i + 3] = ov[i].y;
            }
            if (curvature) {
                curvature[2 * i] = cv[i].x;
                curvature[2 * i + 1] = cv[i].y;
            }
        }
    });
}

fs_status fs_snapshot_save(fs_grid* grid)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(grid->uploaded, FS_E_LOGIC, "Upload the grid state before taking a snapshot");
        set_device(grid->ctx);
        fs_grid& g = *grid;
        auto copy = [&](DevBuf<double>& dst, const DevBuf<double>& src) {
            if (dst.n != src.n)
                g.bytes += dst.alloc(src.n);
            if (src.n)
                HIPC(hipMemcpyAsync(dst.p, src.p, src.n * sizeof(double), hipMemcpyDeviceToDevice, g.st()));
        };
        copy(g.snap_fraction, g.fraction);
        copy(g.snap_u, g.u);
        copy(g.snap_v, g.v);
        copy(g.snap_p, g.p);
        copy(g.snap_p_prev, g.p_prev);
        copy(g.snap_lmass, g.lmass);
        copy(g.snap_gmass, g.gmass);
        HIPC(hipStreamSynchronize(g.st()));
        g.has_snapshot = true;
    });
}

fs_status fs_snapshot_restore(fs_grid* grid)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(grid->has_snapshot, FS_E_LOGIC, "No snapshot to restore");
        set_device(grid->ctx);
        fs_grid& g = *grid;
        // one launch for all seven fields
        CopyList list{};
        auto copy = [&](DevBuf<double>& dst, const DevBuf<double>& src) {
            if (src.n)
                list.item[list.count++] = CopyList::Item{src.p, dst.p, src.n};
        };
        copy(g.fraction, g.snap_fraction);
        copy(g.u, g.snap_u);
        copy(g.v, g.snap_v);
        copy(g.p, g.snap_p);
        copy(g.p_prev, g.snap_p_prev);
        copy(g.lmass, g.snap_lmass);
        copy(g.gmass, g.snap_gmass);
        launch_copy_many(g.st(), list);
        HIPC(hipGetLastError());
        g.coeff_valid = false;
        g.mg_valid = false;
        // no synchronisation: restores are stream-ordered before the next solve
    });
}

fs_status fs_mg_levels(fs_grid* grid, int* levels)
{
    return guard([&] {
        need(grid != nullptr && levels != nullptr, FS_E_INVALID_ARGUMENT, "null pointer");
        *levels = (int)grid->levels.size();
    });
}

fs_status fs_mg_level_shape(fs_grid* grid, int level, int* width, int* height)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(level >= 0 && level < (int)grid->levels.size(), FS_E_INVALID_ARGUMENT, "level out of range");
        if (width) *width = grid->levels[level].w;
        if (height) *height = grid->levels[level].h;
    });
}

fs_status fs_mg_level_centers(fs_grid* grid, int level, double* cx, double* cy)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(level >= 0 && level < (int)grid->levels.size(), FS_E_INVALID_ARGUMENT, "level out of range");
        const Level& L = grid->levels[level];
        set_device(grid->ctx);
        // read back the device copies, which the coarsening kernels use
        if (cx)
            HIPC(hipMemcpy(cx, L.d_cx.p, L.w * sizeof(double), hipMemcpyDeviceToHost));
        if (cy)
            HIPC(hipMemcpy(cy, L.d_cy.p, L.h * sizeof(double), hipMemcpyDeviceToHost));
    });
}

static void ensure_hierarchy(fs_grid& g)
{
    need(g.uploaded, FS_E_LOGIC, "Upload the grid state before inspecting the hierarchy");
    if (!g.coeff_valid)
        g.rebuild_coefficients();
    if (!g.mg_valid) {
        launch_init_scalars(g.st(), g.S.p, 0.0, 0);
        g.rebuild_hierarchy();
        const PcgScalars s = g.read_scalars();
        if (s.mg_failed || s.status == kFailed) {
            g.mg_valid = false;
            fail(FS_E_RUNTIME, "multigrid coarse factorization failed");
        }
    }
}

static std::vector<double> read_padded_interior(const double* dptr, int w, int h)
{
    const int s = w + 2;
    std::vector<double> padded((size_t)s * (h + 2));
    HIPC(hipMemcpy(padded.data(), dptr, padded.size() * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<double> out((size_t)w * h);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            out[(size_t)y * w + x] = padded[(size_t)(y + 1) * s + x + 1];
    return out;
}

fs_status fs_mg_level_weights(fs_grid* grid, int level, double* right, double* down, float* inv_diag)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(level >= 0 && level < (int)grid->levels.size(), FS_E_INVALID_ARGUMENT, "level out of range");
        set_device(grid->ctx);
        ensure_hierarchy(*grid);
        HIPC(hipStreamSynchronize(grid->st()));
        const Level& L = grid->levels[level];
        if (right) {
            auto v = read_padded_interior(L.right64.p, L.w, L.h);
            std::memcpy(right, v.data(), v.size() * sizeof(double));
        }
        if (down) {
            auto v = read_padded_interior(L.down64.p, L.w, L.h);
            std::memcpy(down, v.data(), v.size() * sizeof(double));
        }
        if (inv_diag) {
            const int s = L.w + 2;
            std::vector<float> padded((size_t)s * (L.h + 2));
            HIPC(hipMemcpy(padded.data(), L.invd32.p, padded.size() * sizeof(float), hipMemcpyDeviceToHost));
            for (int y = 0; y < L.h; ++y)
                for (int x = 0; x < L.w; ++x)
                    inv_diag[(size_t)y * L.w + x] = padded[(size_t)(y + 1) * s + x + 1];
        }
    });
}

fs_status fs_mg_level_prolong(fs_grid* grid, int level, int axis, int* first, int* second, double* weight)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(level >= 0 && level + 1 < (int)grid->levels.size(), FS_E_INVALID_ARGUMENT, "level out of range");
        need(axis == 0 || axis == 1, FS_E_INVALID_ARGUMENT, "axis must be 0 or 1");
        const Level& L = grid->levels[level];
        const auto& f = axis == 0 ? L.pxf : L.pyf;
        const auto& s = axis == 0 ? L.pxs : L.pys;
        const auto& w = axis == 0 ? L.pxw : L.pyw;
        for (size_t i = 0; i < f.size(); ++i) {
            if (first) first[i] = f[i];
            if (second) second[i] = s[i];
            if (weight) weight[i] = w[i];
        }
    });
}

fs_status fs_mg_level_restrict(fs_grid* grid, int level, int axis, int* count, int* index, double* weight)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(level >= 0 && level + 1 < (int)grid->levels.size(), FS_E_INVALID_ARGUMENT, "level out of range");
        need(axis == 0 || axis == 1, FS_E_INVALID_ARGUMENT, "axis must be 0 or 1");
        const Level& L = grid->levels[level];
        const auto& c = axis == 0 ? L.rxc : L.ryc;
        const auto& ix = axis == 0 ? L.rxi : L.ryi;
        const auto& w = axis == 0 ? L.rxw : L.ryw;
        for (size_t i = 0; i < c.size(); ++i)
            if (count) count[i] = c[i];
        for (size_t i = 0; i < ix.size(); ++i) {
            if (index) index[i] = ix[i];
            if (weight) weight[i] = w[i];
        }
    });
}

fs_status fs_mg_level_tiles(fs_grid* grid, int level, int* pre, int* post, int* restriction)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(level >= 0 && level < (int)grid->levels.size(), FS_E_INVALID_ARGUMENT, "no such level");
        const Level& L = grid->levels[(size_t)level];
        const bool fused = level < grid->tail_start;
        const bool jacobi = fused && grid->chebyshev_on(level);
        if (pre)
            *pre = fused ? (jacobi ? (L.jacobi_fused >= 0 ? kJacobiVariants + L.jacobi_fused : L.jacobi_pre)
                                   : L.smooth_variant)
                         : -1;
        if (post)
            *post = fused ? (jacobi ? L.jacobi_post : L.post_variant) : -1;
        if (restriction)
            *restriction = fused ? L.restrict_variant : -1;
    });
}

fs_status fs_mg_set_level_tiles(fs_grid* grid, int level, int pre, int post, int restriction)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(level >= 0 && level < grid->tail_start, FS_E_INVALID_ARGUMENT, "no tiled level with this index");
        const bool jacobi = grid->chebyshev_on(level);
        const int smooth_variants = jacobi ? kJacobiVariants : kSmoothVariants;
        Level& L = grid->levels[(size_t)level];
        const Level& Cn = grid->levels[(size_t)level + 1];
        const bool can_fuse = jacobi && 2 * (long long)Cn.w >= L.w && 2 * (long long)Cn.h >= L.h;
        const int pre_variants = smooth_variants + (can_fuse ? kJacobiRestrictVariants : 0);
        need(pre >= -1 && pre < pre_variants && post >= -1 && post < smooth_variants && restriction >= -1 &&
                 restriction < kRestrictVariants,
            FS_E_INVALID_ARGUMENT, "unknown tile variant");
        if (pre >= kJacobiVariants && jacobi)
            L.jacobi_fused = pre - kJacobiVariants;   // the fused pre-smoothing + restriction launch
        else if (pre >= 0 && jacobi) {
            L.jacobi_fused = -1;   // the separate launches with this Jacobi tile
            L.jacobi_pre = pre;
        } else if (pre >= 0)
            L.smooth_variant = pre;
        if (post >= 0)
            (jacobi ? L.jacobi_post : L.post_variant) = post;
        if (restriction >= 0)
            L.restrict_variant = restriction;
        grid->drop_graph();   // the captured iteration holds the old tiles
    });
}

fs_status fs_mg_coarse_factor(fs_grid* grid, double* factor, int* n)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        set_device(grid->ctx);
        ensure_hierarchy(*grid);
        // the bit-exact factorCoarsest factor is formed on demand (the solve uses the sweep inverse)
        const Level& last = grid->levels.back();
        launch_coarse_factor(grid->st(), last.w, last.h, last.right64.p, last.down64.p, grid->factor.p, nullptr,
            grid->coarse_n, nullptr, grid->solid.n > 0);
        HIPC(hipStreamSynchronize(grid->st()));
        if (n)
            *n = grid->coarse_n;
        if (factor && grid->coarse_n > 0)
            HIPC(hipMemcpy(factor, grid->factor.p, (size_t)grid->coarse_n * grid->coarse_n * sizeof(double),
                hipMemcpyDeviceToHost));
    });
}

fs_status fs_mg_apply(fs_grid* grid, const double* residual, double residual_mean, double* correction)
{
    return guard([&] {
        need(grid != nullptr && residual != nullptr && correction != nullptr, FS_E_INVALID_ARGUMENT, "null pointer");
        set_device(grid->ctx);
        fs_grid& g = *grid;
        ensure_hierarchy(g);
        const int W = g.W, H = g.H, s = W + 2;
        std::vector<float> padded(g.NP(), 0.0f);
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x)
                padded[(size_t)(y + 1) * s + x + 1] = (float)residual[(size_t)y * W + x];
        HIPC(hipMemcpy(g.r.p, padded.data(), padded.size() * sizeof(float), hipMemcpyHostToDevice));
        HIPC(hipStreamSynchronize(g.st()));
        PcgScalars init{};
        init.r_mean = residual_mean;
        init.status = kRunning;
        HIPC(hipMemcpy(g.S.p, &init, sizeof(init), hipMemcpyHostToDevice));
        g.vcycle(nullptr);
        HIPC(hipStreamSynchronize(g.st()));
        HIPC(hipMemcpy(padded.data(), g.levels[0].sol0.p, padded.size() * sizeof(float), hipMemcpyDeviceToHost));
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x)
                correction[(size_t)y * W + x] = padded[(size_t)(y + 1) * s + x + 1];
    });
}

fs_status fs_face_weights(fs_grid* grid, double* inv_density_u, double* inv_density_v)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(grid->uploaded, FS_E_LOGIC, "Upload the grid state first");
        set_device(grid->ctx);
        fs_grid& g = *grid;
        if (!g.coeff_valid)
            g.rebuild_coefficients();
        HIPC(hipStreamSynchronize(g.st()));
        const int W = g.W, H = g.H;
        auto right = read_padded_interior(g.levels[0].right64.p, W, H);
        auto down = read_padded_interior(g.levels[0].down64.p, W, H);
        if (inv_density_u) {
            for (int y = 0; y < H; ++y) {
                inv_density_u[(size_t)y * (W + 1)] = 0.0;
                for (int x = 0; x < W; ++x)
                    inv_density_u[(size_t)y * (W + 1) + x + 1] = right[(size_t)y * W + x];
            }
        }
        if (inv_density_v) {
            for (int x = 0; x < W; ++x)
                inv_density_v[x] = 0.0;
            for (int y = 0; y < H; ++y)
                for (int x = 0; x < W; ++x)
                    inv_density_v[(size_t)(y + 1) * W + x] = down[(size_t)y * W + x];
        }
    });
}

// Words of the copy-bandwidth reference: 64 MiB per buffer, twice the aggregate L2, so every launch
// streams from and to HBM.
constexpr size_t kCalibWords = (size_t)16 << 20;

fs_status fs_time_kernel(fs_grid* grid, int kernel, int reps, double* ms_per_launch)
{
    return guard([&] {
        need(grid != nullptr && ms_per_launch != nullptr, FS_E_INVALID_ARGUMENT, "null pointer");
        need(reps > 0, FS_E_INVALID_ARGUMENT, "reps must be positive");
        need(kernel >= 0 && kernel <= 10, FS_E_INVALID_ARGUMENT, "unknown kernel id");
        fs_grid& g = *grid;
        need(g.uploaded, FS_E_LOGIC, "Upload the grid state first");
        set_device(g.ctx);
        if (!g.coeff_valid)
            g.rebuild_coefficients();
        launch_init_scalars(g.st(), g.S.p, 0.0, 1 << 30);
        if ((kernel == 0 || kernel >= 4) && !g.mg_valid)
            g.rebuild_hierarchy();
        LevelView L0 = g.view(0);
        GridDev gd = g.dev();
        const int nlev = (int)g.levels.size();
        const int l1 = std::min(1, g.tail_start);
        if (kernel == 10 && g.calib_src.n != kCalibWords) {
            g.bytes += g.calib_src.alloc(kCalibWords);
            g.bytes += g.calib_dst.alloc(kCalibWords);
        }
        auto launch_once = [&](int i) {
            switch (kernel) {
            case 10: launch_copy_f32(g.st(), g.calib_src.p, g.calib_dst.p, kCalibWords); break;
            case 0:  // the finest post-smoothing launch of a V-cycle (prolongation + four sweeps)
                g.fine_post_smooth(i & 1);
                break;
            case 1: launch_pcg_ad(g.st(), gd); break;
            case 2: launch_pcg_update(g.st(), gd); break;
            case 3: launch_copy_f32(g.st(), g.r.p, g.q.p, g.NP()); break;
            case 4:  // second level's smoother (or the finest, if the tail starts at level 1)
                launch_rb_multi(g.st(), g.view(l1), i & 1, l1 == 0 ? g.r.p : g.view(l1).rhs,
                    l1 == 0 ? g.rmean_ptr() : nullptr, 0, 0, CoarseLink{}, nullptr, l1 == 0,
                    g.levels[l1].smooth_variant);
                break;
            case 5: {
                if (g.tail_start == 0 || nlev < 2)
                    break;
                const Level& F = g.levels[0];
                AxisRestrict rx{F.d_rxc.p, F.d_rxi.p, F.d_rxw.p};
                AxisRestrict ry{F.d_ryc.p, F.d_ryi.p, F.d_ryw.p};
                launch_res_restrict(g.st(), L0, 1, g.r.p, g.rmean_ptr(), g.view(1), rx, ry, nullptr,
                    F.restrict_variant);
                break;
            }
            case 6: launch_mg_tail(g.st(), g.tail, g.tail_lds, g.tail_start == 0 ? g.rmean_ptr() : nullptr,
                        nullptr); break;
            case 7: {
                const Level& last = g.levels.back();
                launch_coarse_inverse(g.st(), last.w, last.h, last.right64.p, last.down64.p, g.ginv.p, g.coarse_n,
                    nullptr, g.solid.n > 0);
                break;
            }
            case 8: launch_pcg_direction(g.st(), gd, g.levels[0].sol0.p); break;
            default: launch_pcg_rz(g.st(), gd, g.levels[0].sol0.p); break;
            }
        };
        launch_once(0);
        HIPC(hipEventRecord(g.ctx->ev0, g.st()));
        for (int i = 0; i < reps; ++i)
            launch_once(i);
        HIPC(hipEventRecord(g.ctx->ev1, g.st()));
        HIPC(hipEventSynchronize(g.ctx->ev1));
        float ms = 0.f;
        HIPC(hipEventElapsedTime(&ms, g.ctx->ev0, g.ctx->ev1));
        *ms_per_launch = (double)ms / reps;
    });
}

fs_status fs_grid_device_bytes(const fs_grid* grid, size_t* bytes)
{
    return guard([&] {
        need(grid != nullptr && bytes != nullptr, FS_E_INVALID_ARGUMENT, "null pointer");
        *bytes = grid->bytes;
    });
}

// The batched projection of the grids listed in `which`, each with its own dt, budget and tolerance: their
// set-up on the context's stream, one persistent launch for those at or under the context's split (the others
// run fs_project's multi-kernel solve beside it), then every velocity correction.  A solo grid's failure
// lands in errors[i].  Returns with the context's stream drained.
static void batch_projection(fs_context* ctx, fs_grid* const* grids, const std::vector<int>& which, const double* dts,
    const int* max_iterations, const double* tolerances, fs_projection_result* results, std::vector<std::string>& errors)
{
    hipStream_t st = ctx->stream;
    const SyncOnExit drain{st};   // the staged descriptors and host results stay valid until then
    // Instances above the split run fs_project's multi-kernel solve on the context's stream while the
    // persistent launch runs the others on a stream of its own; the route depends on the instance's
    // size only, so each result is the same alone and in any batch.
    std::vector<int> batched, solo;
    for (int i : which)
        if (dts[i] > 0.0)
            (batch_solo(ctx, grids[i]->W, grids[i]->H) ? solo : batched).push_back(i);
    const int nb = (int)batched.size();
    const bool split = nb > 0 && !solo.empty();
    if (nb > 0) {
        std::vector<BatchInstance> desc((size_t)nb);
        size_t lds = 0;
        for (int k = 0; k < nb; ++k) {
            const int i = batched[(size_t)k];
            fs_grid& g = *grids[i];
            g.project_begin(dts[i], fs_grid::budget(max_iterations[i]), tolerances[i]);
            desc[(size_t)k] = g.batch_instance();
            if (g.precond == FS_PRECOND_MULTIGRID)
                lds = std::max(lds, g.tail_lds);
        }
        if (ctx->batch_desc.n < (size_t)nb)
            ctx->batch_desc.alloc((size_t)nb);
        // the descriptors go through pinned staging: `desc` dies at the end of this block, before the
        // stream is synchronised
        stage_upload(ctx, ctx->batch_desc.p, desc.data(), desc.size() * sizeof(BatchInstance));
        // One workgroup is one instance's whole solve: while the instances fit on the compute units one
        // each, reserve more than half a unit's 160 KB of LDS so that no two of them share a unit.
        if (nb <= ctx->compute_units && ctx->lds_per_block >= kBatchSpreadLds)
            lds = std::max(lds, kBatchSpreadLds);
        if (lds > 65536)
            need(batch_allow_lds(lds), FS_E_DEVICE, "the device refused the batched solve's LDS request");
        const hipStream_t bs = split ? batch_side_stream(ctx) : st;
        launch_batch_pcg(bs, ctx->batch_desc.p, nb, lds);
        HIPC(hipGetLastError());
        if (split)
            HIPC(hipEventRecord(ctx->batch_out, bs));
    }
    // the solo instances; a failure is kept for its grid while the batched instances are finished
    for (int i : solo) {
        try {
            results[i] = grids[i]->project(dts[i], max_iterations[i], tolerances[i]);
        } catch (const std::exception& e) {
            errors[(size_t)i] = e.what();
        }
    }
    if (nb > 0) {
        if (split)
            HIPC(hipStreamWaitEvent(st, ctx->batch_out, 0));
        for (int i : batched)
            grids[i]->project_end(dts[i]);
        if (ctx->batch_pinned_n < (size_t)nb) {
            std::lock_guard<std::mutex> lock(device_setup_mutex());
            if (ctx->batch_pinned)
                (void)hipHostFree(ctx->batch_pinned);
            ctx->batch_pinned = nullptr;
            ctx->batch_pinned_n = 0;
            HIPC(hipHostMalloc((void**)&ctx->batch_pinned, (size_t)nb * sizeof(PcgScalars), hipHostMallocDefault));
            ctx->batch_pinned_n = (size_t)nb;
        }
        for (int k = 0; k < nb; ++k)
            HIPC(hipMemcpyAsync(ctx->batch_pinned + k, grids[batched[(size_t)k]]->S.p, sizeof(PcgScalars),
                hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));
        for (int k = 0; k < nb; ++k) {
            const int i = batched[(size_t)k];
            const PcgScalars sc = ctx->batch_pinned[k];
            results[i] = grids[i]->project_result(sc, tolerances[i], dts[i]);
            grids[i]->predicted = sc.iterations;
        }
    }
}

fs_status fs_batch_project(int count, fs_grid* const* grids, double dt, int max_iterations,
    const double* divergence_tolerances, fs_projection_result* results)
{
    return guard([&] {
        need(count >= 0 && (count == 0 || (grids && divergence_tolerances && results)), FS_E_INVALID_ARGUMENT,
            "null pointer");
        if (count == 0)
            return;
        fs_context* ctx = batch_context(count, grids, "Upload the grid state before stepping or projecting");
        for (int i = 0; i < count; ++i) {
            need(!grids[i]->active16(), FS_E_UNSUPPORTED,
                "batched solves run FP32 multigrid levels: set the grid's multigrid precision to FS_MG_FP32");
            need(grids[i]->batch_cycle(), FS_E_UNSUPPORTED,
                "batched solves run red-black smoothing: set the grid's smoother to FS_SMOOTHER_RED_BLACK");
        }
        for (int i = 0; i < count; ++i) {
            results[i] = fs_projection_result{};
            results[i].max_divergence = std::numeric_limits<double>::infinity();
            results[i].physical_steps = 1;
        }
        if (dt <= 0.0)
            return;
        set_device(ctx);
        std::vector<int> all((size_t)count);
        for (int i = 0; i < count; ++i)
            all[(size_t)i] = i;
        const std::vector<double> dts((size_t)count, dt);
        const std::vector<int> budgets((size_t)count, max_iterations);
        std::vector<std::string> errors((size_t)count);
        batch_projection(ctx, grids, all, dts.data(), budgets.data(), divergence_tolerances, results, errors);
        for (const std::string& e : errors)
            if (!e.empty())
                fail(FS_E_RUNTIME, e);
    });
}

fs_status fs_batch_step(int count, fs_grid* const* grids, const double* dts, const int* max_iterations,
    const double* divergence_tolerances, fs_projection_result* results, fs_status* statuses)
{
    return guard([&] {
        need(count >= 0 && (count == 0 || (grids && dts && max_iterations && divergence_tolerances && results)),
            FS_E_INVALID_ARGUMENT, "null pointer");
        if (count == 0)
            return;
        fs_context* ctx = batch_context(count, grids, "Upload the grid state before stepping");
        const size_t n = (size_t)count;
        // fs_step's plan per grid: `count` capillary substeps of dt / count (FluidGridCPU.cpp:319-357)
        std::vector<int> substeps(n, 0);
        std::vector<double> step(n, 0.0);
        std::vector<std::string> errors(n);
        std::vector<char> alive(n, 0);
        for (size_t i = 0; i < n; ++i) {
            fs_projection_result& r = results[i];
            r = fs_projection_result{};
            r.max_divergence = std::numeric_limits<double>::infinity();
            if (statuses)
                statuses[i] = FS_OK;
            const fs_grid& g = *grids[i];
            if (!(dts[i] > 0.0))   // FluidGridCPU::step returns {0, inf, false, 0, 0} for dt <= 0
                continue;
            const double subdivisions = g.sigma > 0.0 ? std::max(1.0, std::ceil(dts[i] / g.capillary_step_limit())) : 1.0;
            if (subdivisions > (double)std::numeric_limits<int>::max()) {
                errors[i] = "Surface tension requires too many physical substeps for this timestep";
                continue;
            }
            substeps[i] = (int)subdivisions;
            step[i] = dts[i] / substeps[i];
            alive[i] = 1;
        }
        set_device(ctx);
        std::vector<int> viscosity_iterations(n, 0);
        std::vector<fs_projection_result> proj(n);
        for (int k = 0;; ++k) {
            std::vector<int> active;
            for (size_t i = 0; i < n; ++i)
                if (alive[i] && k < substeps[i])
                    active.push_back((int)i);
            if (active.empty())
                break;
            auto keep = [&](std::vector<int>& list) {
                list.erase(std::remove_if(list.begin(), list.end(), [&](int i) { return !alive[(size_t)i]; }),
                    list.end());
            };
            auto attempt = [&](int i, auto&& body) {
                try {
                    body();
                } catch (const std::exception& e) {
                    errors[(size_t)i] = e.what();
                    alive[(size_t)i] = 0;
                }
            };
            // transport, one grid after another (each reads its CFL bound back)
            for (int i : active)
                attempt(i, [&] {
                    fs_grid& g = *grids[i];
                    if (g.transport == FS_TRANSPORT_REFERENCE) {
                        advect_reference(g, step[(size_t)i]);
                    } else if (g.transport == FS_TRANSPORT_SEMI_LAGRANGIAN) {
                        advect_sl(g, step[(size_t)i]);
                        results[i].transport_steps += 1;
                    }
                });
            keep(active);
            // every viscous grid's implicit solve, batched
            std::fill(viscosity_iterations.begin(), viscosity_iterations.end(), 0);
            batch_viscosity(ctx, grids, active, step.data(), viscosity_iterations.data(), errors);
            for (int i : active) {
                if (!errors[(size_t)i].empty())
                    alive[(size_t)i] = 0;
                results[i].viscosity_iterations += viscosity_iterations[(size_t)i];
            }
            keep(active);
            // body forces
            for (int i : active)
                attempt(i, [&] {
                    fs_grid& g = *grids[i];
                    launch_gravity(g.st(), g.dev(), g.gravity_y, step[(size_t)i]);
                    if (g.sigma > 0.0)
                        g.apply_surface_tension(step[(size_t)i]);
                });
            keep(active);
            // the projections: FP32-level grids with the default cycle in one batched solve, the others one by one
            std::vector<int> batched, alone;
            for (int i : active)
                (grids[i]->active16() || !grids[i]->batch_cycle() ? alone : batched).push_back(i);
            batch_projection(ctx, grids, batched, step.data(), max_iterations, divergence_tolerances, proj.data(), errors);
            for (int i : alone)
                attempt(i, [&] {
                    proj[(size_t)i] = grids[i]->project(step[(size_t)i], max_iterations[i], divergence_tolerances[i]);
                });
            for (int i : active) {
                if (!errors[(size_t)i].empty()) {
                    alive[(size_t)i] = 0;
                    continue;
                }
                const fs_projection_result& p = proj[(size_t)i];
                fs_projection_result& r = results[i];
                r.iterations += p.iterations;
                r.max_divergence = p.max_divergence;
                r.converged = p.converged;
                r.restarts = p.restarts;
                r.initial_residual = p.initial_residual;
                r.final_residual = p.final_residual;
                r.rhs_max = p.rhs_max;
                r.physical_steps += 1;
                if (!r.converged)
                    alive[(size_t)i] = 0;   // stop after the first substep that does not converge
            }
        }
        std::string first;
        for (size_t i = 0; i < n; ++i) {
            if (errors[i].empty())
                continue;
            if (statuses)
                statuses[i] = FS_E_RUNTIME;
            if (first.empty())
                first = errors[i];
        }
        if (!first.empty())
            fail(FS_E_RUNTIME, first);
    });
}

fs_status fs_tune_batch_smoother(int count, fs_grid* const* grids, double dt, int max_iterations,
    const double* divergence_tolerances, int* sweeps, int* coarse_sweeps, double* ms)
{
    return guard([&] {
        need(count > 0 && grids && divergence_tolerances, FS_E_INVALID_ARGUMENT, "null pointer or empty batch");
        need(std::isfinite(dt) && dt > 0.0, FS_E_INVALID_ARGUMENT, "dt must be finite and positive");
        fs_context* ctx = batch_context(count, grids, "Upload the grid state before tuning");
        for (int i = 0; i < count; ++i)
            need(!grids[i]->active16(), FS_E_UNSUPPORTED,
                "batched solves run FP32 multigrid levels: set the grid's multigrid precision to FS_MG_FP32");
        set_device(ctx);
        const size_t n = (size_t)count;
        std::vector<ProjectionState> saved(n);
        std::vector<SmootherChoice> original(n);
        for (size_t i = 0; i < n; ++i) {
            fs_grid& g = *grids[i];
            g.drain();
            saved[i].save(g);
            original[i] = {g.smoother, g.smooth_sweeps, g.coarse_sweeps};
        }
        HIPC(hipStreamSynchronize(ctx->stream));
        auto choose_all = [&](const SmootherChoice& c) {
            HIPC(hipStreamSynchronize(ctx->stream));
            for (size_t i = 0; i < n; ++i)
                grids[i]->set_smoother(c.kind, c.sweeps, c.coarse);
        };
        auto restore_all = [&] {
            for (size_t i = 0; i < n; ++i)
                saved[i].restore(*grids[i]);
            HIPC(hipStreamSynchronize(ctx->stream));
        };
        std::vector<int> all(n);
        for (size_t i = 0; i < n; ++i)
            all[i] = (int)i;
        const std::vector<double> dts(n, dt);
        const std::vector<int> budgets(n, max_iterations);
        std::vector<fs_projection_result> results(n);
        std::vector<std::string> errors(n);
        SmootherChoice best{FS_SMOOTHER_RED_BLACK, 4, 4};
        double best_ms = std::numeric_limits<double>::infinity();
        try {
            for (int coarse : {4, 2})
                for (int nu : {4, 3, 2}) {
                    const SmootherChoice c{FS_SMOOTHER_RED_BLACK, nu, coarse};
                    choose_all(c);
                    double fastest = std::numeric_limits<double>::infinity();
                    bool ok = true;
                    for (int rep = 0; rep < 2; ++rep) {   // the first is a warm-up
                        restore_all();
                        const auto t0 = std::chrono::steady_clock::now();
                        batch_projection(ctx, grids, all, dts.data(), budgets.data(), divergence_tolerances,
                            results.data(), errors);
                        const double t =
                            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                        for (size_t i = 0; i < n; ++i)
                            ok = ok && errors[i].empty() && results[i].converged;
                        if (rep > 0)
                            fastest = t;
                    }
                    if (ok && fastest < best_ms) {
                        best_ms = fastest;
                        best = c;
                    }
                }
        } catch (...) {
            HIPC(hipStreamSynchronize(ctx->stream));
            for (size_t i = 0; i < n; ++i)
                grids[i]->set_smoother(original[i].kind, original[i].sweeps, original[i].coarse);
            restore_all();
            throw;
        }
        choose_all(best);
        restore_all();
        if (sweeps)
            *sweeps = best.sweeps;
        if (coarse_sweeps)
            *coarse_sweeps = best.coarse;
        if (ms)
            *ms = best_ms;
    });
}

fs_status fs_context_set_batch_split(fs_context* ctx, long long solo_cells)
{
    return guard([&] {
        need(ctx != nullptr, FS_E_INVALID_ARGUMENT, "null context");
        need(solo_cells >= 0, FS_E_INVALID_ARGUMENT, "the split must be a non-negative cell count");
        ctx->batch_solo_cells = solo_cells;
    });
}

fs_status fs_grid_estimate_bytes(int width, int height, int n_liquid, int n_gas, int features, size_t* bytes)
{
    return guard([&] {
        need(bytes != nullptr, FS_E_INVALID_ARGUMENT, "null pointer");
        need(width > 0 && height > 0 && n_liquid >= 0 && n_gas >= 0, FS_E_INVALID_ARGUMENT,
            "grid dimensions must be positive and component counts non-negative");
        const size_t W = (size_t)width, H = (size_t)height, N = W * H, NU = (W + 1) * H, NV = W * (H + 1);
        const size_t NP = (W + 2) * (H + 2), nl = (size_t)n_liquid, ng = (size_t)n_gas;
        const size_t blocks = ((W + 63) / 64) * ((H + 15) / 16);
        // state, pressure history, PCG vectors and reductions (fs_grid::allocate)
        size_t b = 8 * (3 * N + NU + NV + 3 * NP) + 4 * 5 * NP + 8 * (6 * blocks + 64) +
                   4 * sizeof(PcgScalars) + 8 * (nl + ng) * N;
        // pressure hierarchy (build_hierarchy): FP64 + FP32 conductances, 1/diag, rhs, two solutions,
        // residual, transfer tables; dense coarse factor and inverse; tail tables
        size_t w = W, h = H;
        for (int l = 0;; ++l) {
            const size_t np = (w + 2) * (h + 2);
            b += np * (16 + 12 + 12 + (l > 0 ? 4 : 0)) + 64 * (w + h);
            if (w * h <= 64)
                break;
            w = (w + 1) / 2;
            h = (h + 1) / 2;
        }
        const size_t coarse = std::max<size_t>(w * h, 2) - 1;
        b += coarse * coarse * 12 + (64 << 10);
        if (features & FS_FEATURE_TRANSPORT)   // advect_reference buffers
            b += 8 * (N + 2 * (nl + ng) * N + 3 * NU + 3 * NV + 2 * std::max(NU, NV) + N) + N + 8;
        if (features & FS_FEATURE_VISCOSITY) {   // ViscositySolver + blended material
            size_t vw = W, vh = H;
            for (;;) {
                const size_t ns = (vw + 1) * (vh + 1);
                b += ns * (16 + 4 * 8) + 8 * vw * vh;
                if (vw * vh <= 64)
                    break;
                vw = (vw + 1) / 2;
                vh = (vh + 1) / 2;
            }
            b += 6 * 8 * (W + 1) * (H + 1) + 2 * 16 * N + 32 * blocks + 4096;
        }
        if (features & FS_FEATURE_SURFACE_TENSION)
            b += 4 * 16 * N;
        if (features & FS_FEATURE_SOLID_MASK)
            b += N;
        if (features & FS_FEATURE_MG16) {   // 16-bit conductances and (bfloat16 vectors) rhs + 2 solutions
            size_t lw = W, lh = H;
            for (;;) {
                b += (lw + 2) * (lh + 2) * 2 * 5;
                if (lw * lh <= 64)
                    break;
                lw = (lw + 1) / 2;
                lh = (lh + 1) / 2;
            }
        }
        *bytes = b;
    });
}

fs_status fs_get_surface_tension(const fs_grid* grid, double* coefficient)
{
    return guard([&] {
        need(grid != nullptr && coefficient != nullptr, FS_E_INVALID_ARGUMENT, "null pointer");
        *coefficient = grid->sigma;
    });
}

fs_status fs_get_preconditioner(const fs_grid* grid, fs_preconditioner* kind)
{
    return guard([&] {
        need(grid != nullptr && kind != nullptr, FS_E_INVALID_ARGUMENT, "null pointer");
        *kind = grid->precond;
    });
}

fs_status fs_get_transport(const fs_grid* grid, fs_transport* mode)
{
    return guard([&] {
        need(grid != nullptr && mode != nullptr, FS_E_INVALID_ARGUMENT, "null pointer");
        *mode = grid->transport;
    });
}

fs_status fs_component_counts(const fs_grid* grid, int* n_liquid, int* n_gas)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        if (n_liquid) *n_liquid = (int)grid->rho_l.size();
        if (n_gas) *n_gas = (int)grid->rho_g.size();
    });
}

fs_status fs_get_component(const fs_grid* grid, int phase, int component, double* density, double* viscosity)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(phase == 0 || phase == 1, FS_E_INVALID_ARGUMENT, "phase must be 0 (liquid) or 1 (gas)");
        const std::vector<double>& rho = phase == 0 ? grid->rho_l : grid->rho_g;
        const std::vector<double>& mu = phase == 0 ? grid->mu_l : grid->mu_g;
        need(component >= 0 && (size_t)component < rho.size(), FS_E_INVALID_ARGUMENT, "Material component is out of range");
        if (density) *density = rho[component];
        if (viscosity) *viscosity = mu[component];
    });
}

fs_status fs_checkpoint_size(const fs_grid* grid, size_t* metadata_bytes, size_t* payload_bytes)
{
    return guard([&] {
        need(grid != nullptr, FS_E_INVALID_ARGUMENT, "null grid");
        need(grid->uploaded, FS_E_LOGIC, "Upload before saving a checkpoint");
        if (metadata_bytes)
            *metadata_bytes = checkpoint_metadata(*grid).size() + 1;
        if (payload_bytes)
            *payload_bytes = checkpoint_values(*grid, grid->rho_l.size(), grid->rho_g.size()) * sizeof(double);
    });
}

fs_status fs_checkpoint_save(fs_grid* grid, char* metadata, size_t metadata_capacity, void* payload,
    size_t payload_capacity)
{
    return guard([&] {
        need(grid != nullptr && metadata != nullptr && payload != nullptr, FS_E_INVALID_ARGUMENT, "null pointer");
        need(grid->uploaded, FS_E_LOGIC, "Upload before saving a checkpoint");
        fs_grid& g = *grid;
        const std::string text = checkpoint_metadata(g);
        const size_t values = checkpoint_values(g, g.rho_l.size(), g.rho_g.size());
        need(metadata_capacity >= text.size() + 1 && payload_capacity >= values * sizeof(double),
            FS_E_INVALID_ARGUMENT, "checkpoint buffers are smaller than fs_checkpoint_size reports");
        set_device(g.ctx);
        unsigned char* out = static_cast<unsigned char*>(payload);
        size_t offset = 0;
        auto get = [&](const double* src, size_t n) {
            if (n)
                HIPC(hipMemcpyAsync(out + offset, src, n * sizeof(double), hipMemcpyDeviceToHost, g.st()));
            offset += n * sizeof(double);
        };
        const size_t N = g.N();
        get(g.fraction.p, N);
        get(g.u.p, g.NU());
        get(g.v.p, g.NV());
        // Both pressure histories are stored without the halo; the staging buffer is reused in
        // stream order.
        launch_pad_copy(g.st(), g.p.p, g.tmp64.p, g.W, g.H, 0);
        get(g.tmp64.p, N);
        launch_pad_copy(g.st(), g.p_prev.p, g.tmp64.p, g.W, g.H, 0);
        get(g.tmp64.p, N);
        get(g.lmass.p, g.rho_l.size() * N);
        get(g.gmass.p, g.rho_g.size() * N);
        HIPC(hipStreamSynchronize(g.st()));
        std::memcpy(metadata, text.c_str(), text.size() + 1);
    });
}

fs_status fs_checkpoint_load(fs_grid* grid, const char* metadata, size_t metadata_length, const void* payload,
    size_t payload_bytes)
{
    return guard([&] {
        need(grid != nullptr && metadata != nullptr, FS_E_INVALID_ARGUMENT, "null pointer");
        fs_grid& g = *grid;
        const CheckpointConfig c = parse_checkpoint(g, metadata, metadata_length);
        const size_t values = checkpoint_values(g, c.rho_l.size(), c.rho_g.size());
        if (payload_bytes != values * sizeof(double) || (values && payload == nullptr))
            bad_checkpoint("fluid payload has the wrong size");
        // Validated: replace the live configuration and state.
        set_device(g.ctx);
        g.drain();
        const size_t N = g.N();
        if (g.lmass.n != c.rho_l.size() * N)
            g.bytes += g.lmass.alloc(c.rho_l.size() * N);
        if (g.gmass.n != c.rho_g.size() * N)
            g.bytes += g.gmass.alloc(c.rho_g.size() * N);
        g.rho_l = c.rho_l;
        g.rho_g = c.rho_g;
        g.mu_l = c.mu_l;
        g.mu_g = c.mu_g;
        g.ox = c.ox;
        g.oy = c.oy;
        g.gravity_y = c.gravity;
        g.sigma = c.sigma;
        g.precond = c.precond;
        g.transport = c.transport;
        g.predicted = c.predicted;
        const unsigned char* in = static_cast<const unsigned char*>(payload);
        size_t offset = 0;
        auto put = [&](double* dst, size_t n) {
            if (n)
                HIPC(hipMemcpyAsync(dst, in + offset, n * sizeof(double), hipMemcpyHostToDevice, g.st()));
            offset += n * sizeof(double);
        };
        put(g.fraction.p, N);
        put(g.u.p, g.NU());
        put(g.v.p, g.NV());
        put(g.tmp64.p, N);
        launch_pad_copy(g.st(), g.tmp64.p, g.p.p, g.W, g.H, 1);
        put(g.tmp64.p, N);
        launch_pad_copy(g.st(), g.tmp64.p, g.p_prev.p, g.W, g.H, 1);
        put(g.lmass.p, c.rho_l.size() * N);
        put(g.gmass.p, c.rho_g.size() * N);
        HIPC(hipStreamSynchronize(g.st()));
        g.uploaded = true;
        g.coeff_valid = false;
        g.mg_valid = false;
        g.has_snapshot = false;
    });
}

fs_status fs_advance_frame(fs_grid* grid, const fs_frame_settings* settings, fs_frame_statistics* statistics)
{
    fs_frame_statistics stats{};
    const fs_status status = guard([&] {
        need(grid != nullptr && settings != nullptr, FS_E_INVALID_ARGUMENT, "null pointer");
        const fs_frame_settings& s = *settings;
        need(std::isfinite(s.frame_time) && s.frame_time > 0.0, FS_E_INVALID_ARGUMENT,
            "frame time must be finite and positive");
        need(s.steps_per_second > 0, FS_E_INVALID_ARGUMENT, "steps per second must be positive");
        need(grid->uploaded, FS_E_LOGIC, "Upload the grid state before stepping");
        stats.projection_tolerance = s.projection_tolerance;
        // FluidScene::sceneProcess (FluidScene.inl:343-362): equal substeps at the scene's minimum rate,
        // or one engine step when the capillary subdivision already reaches that rate; uniform GPU
        // scenes solve to tolerance without an iteration cutoff.
        const double wanted = std::ceil(s.frame_time * (double)s.steps_per_second);
        need(wanted <= 1e6, FS_E_INVALID_ARGUMENT, "frame needs more than 1e6 steps");
        int step_count = (int)wanted;
        if (std::ceil(s.frame_time / grid->capillary_step_limit()) >= step_count)
            step_count = 1;
        const double step = s.frame_time / step_count;
        stats.steps_scheduled = step_count;
        stats.accepted = 1;
        using Clock = std::chrono::steady_clock;
        for (int k = 0; k < step_count; ++k) {
            fs_step_timings t{};
            fs_projection_result r{};
            const auto started = Clock::now();
            const fs_status st = fs_step(grid, step, -1, s.projection_tolerance, s.collect_timings ? &t : nullptr, &r);
            stats.fluid_ms += std::chrono::duration<double, std::milli>(Clock::now() - started).count();
            if (st != FS_OK) {
                stats.accepted = 0;
                fail(st, g_last_error);
            }
            stats.stages.transport_ms += t.transport_ms;
            stats.stages.viscosity_ms += t.viscosity_ms;
            stats.stages.forces_ms += t.forces_ms;
            stats.stages.pressure_ms += t.pressure_ms;
            stats.stages.selection_ms += t.selection_ms;
            stats.stages.regrid_ms += t.regrid_ms;
            stats.substeps += r.physical_steps;
            stats.transport_steps += r.transport_steps;
            stats.pressure_iterations += r.iterations;
            stats.viscosity_iterations += r.viscosity_iterations;
            stats.max_volume_error = std::max(stats.max_volume_error, r.max_volume_error);
            stats.rms_volume_error = std::max(stats.rms_volume_error, r.rms_volume_error);
            stats.max_divergence = r.max_divergence;
            ++stats.steps_run;
            // a projection that is not accepted ends the frame (FluidScene.inl:387-393)
            if (!(r.converged || r.relaxed || r.precision_limited)) {
                stats.accepted = 0;
                break;
            }
        }
    });
    if (statistics)
        *statistics = stats;
    return status;
}

}  // extern "C"
