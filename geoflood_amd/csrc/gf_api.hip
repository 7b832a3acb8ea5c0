// gf_api.hip -- host runtime and C ABI of libgeoflood_amd.so.
//
// Owns the HIP stream, the device-resident patch pools, the pinned/device
// staging used by the drop-in batch entry point, the hipGraph of one step
// attempt, and the HIP-event timing of the fused step launches.  No HIP
// allocation happens per step: pools and staging are sized at creation (the
// drop-in staging grows only when a larger batch arrives, as buffer-len does
// in the reference).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <new>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/geoflood_amd.h"
#include "gf_launch.h"
#include "gf_output.h"

using gf::PatchDesc;
using gf::ScenCtrl;

// Device memory of gf_step2_batch_noalloc: a fixed array per GPU, loaded
// with the library's code object and held until the process ends.  Host code
// reaches it only through the address hipGetSymbolAddress returns on each
// device (Lib::scratch_dev); the name itself is not a device address there.
constexpr size_t kNoAllocBytes = size_t(32) << 20;
static __device__ __align__(256) double scratch[kNoAllocBytes / sizeof(double)];

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

// Every entry point of the C ABI is a function-try-block: no C++ exception
// (a failed host allocation, a thread that cannot start) crosses into a C
// caller; it becomes a status and a message.
#define GF_CATCH_STATUS                                                           \
    catch (const std::bad_alloc&) { return fail(GF_ENOMEM, "host memory allocation failed"); } \
    catch (const std::exception& e_) { return fail(GF_ESTATE, "host error: %s", e_.what()); } \
    catch (...) { return fail(GF_ESTATE, "host error"); }
#define GF_CATCH_NULL                                                             \
    catch (const std::bad_alloc&) { fail(GF_ENOMEM, "host memory allocation failed"); return nullptr; } \
    catch (const std::exception& e_) { fail(GF_ESTATE, "host error: %s", e_.what()); return nullptr; } \
    catch (...) { fail(GF_ESTATE, "host error"); return nullptr; }

#define HIPCHK(x)                                                                       \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
        if (e_ != hipSuccess) return fail(GF_EHIP, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

// Host threads kept for the staging copies between pageable patch memory and
// the pinned staging buffers: run(n, f) calls f(0..n-1) over the workers and
// the calling thread and returns when every call has ended.
class HostWorkers {
public:
    explicit HostWorkers(int nworkers)
    {
        try {
            for (int t = 0; t < nworkers; t++) th_.emplace_back([this] { loop(); });
        } catch (...) {
            stop();  // the threads already started end before the exception leaves
            throw;
        }
    }
    ~HostWorkers() { stop(); }
    void run(int n, const std::function<void(int)>& f)
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = &f;
            n_ = n;
            done_ = 0;
            next_.store(0);
            gen_++;
        }
        go_.notify_all();
        work();
        std::unique_lock<std::mutex> lk(mu_);
        // every worker reports each generation, so none can lag into the next
        finished_.wait(lk, [this] { return done_ == (int)th_.size(); });
        job_ = nullptr;
    }

private:
    void stop()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            quit_ = true;
        }
        go_.notify_all();
        for (auto& t : th_) t.join();
        th_.clear();
    }
    void work()
    {
        for (int k; (k = next_.fetch_add(1)) < n_;) (*job_)(k);
    }
    void loop()
    {
        unsigned seen = 0;
        for (;;) {
            std::unique_lock<std::mutex> lk(mu_);
            go_.wait(lk, [&] { return quit_ || gen_ != seen; });
            if (quit_) return;
            seen = gen_;
            lk.unlock();
            work();
            lk.lock();
            if (++done_ == (int)th_.size()) finished_.notify_one();
        }
    }
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable go_, finished_;
    const std::function<void(int)>* job_ = nullptr;
    int n_ = 0, done_ = 0;
    std::atomic<int> next_{0};
    unsigned gen_ = 0;
    bool quit_ = false;
};

// The solver's constants of grav and dry_tolerance (gf::Params::halfg ...),
// each formed exactly as the per-face expression it replaces.
static void derive_params(gf::Params& P)
{
    P.halfg = 0.5 * P.grav;
    const double c = P.dry_tolerance * P.grav;
    P.crit = c > 1e-6 ? c : 1e-6;
    P.crit2 = std::sqrt(P.crit);
    P.critsq = P.crit * P.crit;
    P.inv16g = 1.0 / (16.0 * P.grav);
}

struct Lib {
    bool init = false;
    int device = 0;
    hipStream_t stream = nullptr;
    gf::Params P{};
    gf_params user{};
    int num_cu = 256;
    hipEvent_t ev_begin = nullptr, ev_end = nullptr;
    hipEvent_t ev_half[2] = {nullptr, nullptr};  // the two halves of a pool's staging buffers
    HostWorkers* workers = nullptr;  // created on first use, joined by gf_finalize
    std::vector<hipEvent_t> ev_step;  // pairs around step launches
    std::vector<hipEvent_t> ev_tune;  // pair around the autotuner's candidate launches
    unsigned params_gen = 1;  // bumped by gf_set_params: pools re-sync their batch arguments
    double last_total_ms = 0.0, last_step_ms = 0.0;
    int last_launches = 0;
    // drop-in batch staging
    gf_pool* batch_pool = nullptr;
    double* h_stage = nullptr;  // pinned
    size_t h_stage_bytes = 0;
    double* d_max = nullptr;
    double* h_max = nullptr;  // pinned
    // per device: the address of `scratch` there (set by gf_init)
    static constexpr int kMaxDevices = 64;
    double* scratch_dev[kMaxDevices] = {};
};
Lib g;

constexpr size_t kChunk = 1024;  // patches per upload/download chunk

// Resident AMR: refinement levels 0 .. kMaxLevels-1; device ranges
// [lo, hi) of the stored links (all, then the interpolation links of each
// coarse level) and of the inflow list's entries per level.
constexpr int kMaxLevels = 16;
constexpr int kRanges = 2 * (1 + 2 * kMaxLevels);
inline int range_links() { return 0; }
inline int range_interp(int l) { return 2 + 2 * l; }
inline int range_inflow(int l) { return 2 + 2 * kMaxLevels + 2 * l; }
constexpr int kLinkGrid = 1024;  // workgroups of a resident coarse/fine transfer (one wave per link, grid-stride)

// Maximum in which a NaN operand wins, as the device reductions fold CFL
// numbers: a patch whose new state is not finite reports a NaN CFL number.
double max_nan(double a, double b) { return (b > a || b != b) ? b : a; }

// Host memory that the copy engines read directly (hipHostMalloc), for the
// pool's descriptor table: a regrid re-describes every leaf (72 bytes each,
// 1.7 MB at 23k leaves), and an upload from pageable memory goes through
// the runtime's staging copy first.  Every upload from it is followed by a
// stream synchronisation before the table is written again.
template <typename T>
struct PinnedAlloc {
    using value_type = T;
    PinnedAlloc() = default;
    template <typename U>
    PinnedAlloc(const PinnedAlloc<U>&) {}
    T* allocate(size_t n)
    {
        void* m = nullptr;
        if (hipHostMalloc(&m, n * sizeof(T), hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            throw std::bad_alloc();
        }
        return static_cast<T*>(m);
    }
    void deallocate(T* m, size_t) { (void)hipHostFree(m); }
};
template <typename T, typename U>
bool operator==(const PinnedAlloc<T>&, const PinnedAlloc<U>&) { return true; }
template <typename T, typename U>
bool operator!=(const PinnedAlloc<T>&, const PinnedAlloc<U>&) { return false; }

}  // namespace

struct gf_pool {
    int mx = 0, my = 0, mbc = 2, capacity = 0, n = 0;
    size_t np2 = 0;
    double* q[2] = {nullptr, nullptr};
    double* aux = nullptr;
    PatchDesc* desc = nullptr;
    double* cfl = nullptr;
    double* hmax = nullptr;
    ScenCtrl* ctrl = nullptr;
    double* d_stage = nullptr;  // device staging (kChunk patches of q and aux)
    double* h_stage = nullptr;  // pinned host staging
    size_t stage_doubles = 0;   // size of each staging buffer
    double* h_cfl = nullptr;    // pinned
    // batched path
    double* hmax2 = nullptr;               // [2][capacity]
    unsigned long long* cflbits = nullptr; // [capacity] per scenario
    int* list_wet = nullptr;
    int* list_dry = nullptr;
    int* counters = nullptr;               // gf::NCOUNTERS
    long long* stats = nullptr;            // [4] see gf_pool_stats
    gf::BatchArgs* d_args = nullptr;       // device copy read by the persistent kernel
    unsigned long long* tile_cfl = nullptr; // [capacity] tile fold slots of tiled step variants
    unsigned long long* tile_h = nullptr;
    int* tile_cnt = nullptr;
    int* dry_age = nullptr;                // [capacity] steady dry patches (gf::DRY_STEADY)
    int variant = 0;                       // step variant (tile shape, threads); see gf_pool_autotune
    int wet_grid = 0;                      // workgroups of the persistent step kernel (0: not computed)
    int* inflow_list = nullptr;            // [capacity] patches with a hydrograph side
    int* inflow_n = nullptr;               // [1] entries of inflow_list
    bool has_inflow = false;               // inflow_list is not empty (the graph launches k_inflow)
    std::vector<double*> hyd_mem;          // per scenario slot: device hydrograph table
    std::string inflow_error;              // set when a scenario's inflow cannot run
    void* tag_buf = nullptr;               // criteria, patch list and tags of the tagging calls
    size_t tag_bytes = 0;
    bool args_dirty = true;
    unsigned args_gen = 0;                 // params generation of d_args and of the graph
    int args_n = -1, args_nscen = -1;      // the slot and scenario counts in d_args
    std::vector<PatchDesc, PinnedAlloc<PatchDesc>> h_desc;
    std::vector<ScenCtrl> h_ctrl;
    int nscen = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    bool graph_dirty = true;
    // Resident AMR (gf_pool_set_levels, gf_pool_set_cf_links): every attempt
    // starts with the multi-level ghost update -- fine -> coarse averaging
    // over all links, then level by level the hydrograph sides and the
    // coarse -> fine interpolation -- whose link and inflow ranges live in
    // device memory, so a regrid changes them without a new graph.  The
    // graph holds only the passes whose ranges are not empty (a forest that
    // occupies levels 3-4 launches no pass for levels 0-2); it is captured
    // again when that set changes, which a regrid that keeps the occupied
    // levels does not.
    bool amr = false;
    std::vector<int> h_level;              // [capacity] refinement level per slot (0 when not set)
    int nlev = 1;                          // levels the ghost update walks (max level + 1)
    int* d_links = nullptr;                // [4][link_cap]: coarse, fine, side, igrid by the coarse level
    int link_cap = 0, nlinks = 0;
    unsigned inflow_levels = 0;            // bit l: level l has hydrograph patches (its k_inflow pass runs)
    unsigned interp_levels = 0;            // bit l: links with a level-l coarse patch (its interpolation runs)
    int* d_ranges = nullptr;               // [kRanges]: see range_links / range_interp / range_inflow
};

static void default_params()
{
    gf_params p{};
    p.grav = 9.81;
    p.dry_tolerance = 1e-3;
    p.manning_n = 0.025;
    p.friction_depth = 1.0e6;
    p.friction_forcing = 1;
    p.src_term = 1;
    p.order = 2;
    p.transverse = 2;
    p.mthlim[0] = p.mthlim[1] = p.mthlim[2] = 4;
    p.mcapa = 0;
    g.user = p;
    g.P.grav = p.grav;
    g.P.dry_tolerance = p.dry_tolerance;
    g.P.manning_n = p.manning_n;
    g.P.friction_depth = p.friction_depth;
    g.P.friction_forcing = p.friction_forcing;
    g.P.src_term = p.src_term;
    g.P.order = p.order;
    g.P.transverse = p.transverse;
    for (int k = 0; k < 3; k++) g.P.mthlim[k] = p.mthlim[k];
    derive_params(g.P);
}

static int need_init()
{
    if (!g.init) return fail(GF_ESTATE, "gf_init has not been called");
    return GF_OK;
}

static int ensure_step_events(size_t pairs)
{
    while (g.ev_step.size() < 2 * pairs) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        g.ev_step.push_back(e);
    }
    return GF_OK;
}

static gf::StepArgs step_args(gf_pool* p, int first, int count, double* q0, double* q1,
                              ScenCtrl* ctrl, double dt, int cur)
{
    gf::StepArgs a{};
    a.desc = p->desc;
    a.list = nullptr;
    a.first = first;
    a.count = count;
    a.q0 = q0;
    a.q1 = q1;
    a.aux = p->aux;
    a.cfl_out = p->cfl;
    a.hmax_out = p->hmax2 ? p->hmax2 + (size_t)(q1 ? (cur ^ 1) : cur) * p->capacity : nullptr;
    a.ctrl = ctrl;
    a.dt = dt;
    a.cur = cur;
    a.P = g.P;
    a.tile_cfl = p->tile_cfl;
    a.tile_h = p->tile_h;
    a.tile_cnt = p->tile_cnt;
    return a;
}

extern "C" {

int gf_init(int device)
try {
    int count = 0;
    HIPCHK(hipGetDeviceCount(&count));
    if (count <= 0) return fail(GF_EHIP, "no HIP device");
    if (device < 0 || device >= count) return fail(GF_EINVAL, "device %d out of range [0,%d)", device, count);
    if (g.init) {
        if (device == g.device) return GF_OK;
        return fail(GF_ESTATE, "already initialised on device %d", g.device);
    }
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&g.ev_begin));
    HIPCHK(hipEventCreate(&g.ev_end));
    HIPCHK(hipEventCreateWithFlags(&g.ev_half[0], hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&g.ev_half[1], hipEventDisableTiming));
    HIPCHK(hipDeviceGetAttribute(&g.num_cu, hipDeviceAttributeMultiprocessorCount, device));
    HIPCHK(hipMalloc(&g.d_max, sizeof(double)));
    HIPCHK(hipHostMalloc(&g.h_max, sizeof(double), hipHostMallocDefault));
    if (device < Lib::kMaxDevices && !g.scratch_dev[device]) {
        void* sp = nullptr;
        HIPCHK(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
        g.scratch_dev[device] = static_cast<double*>(sp);
    }
    g.device = device;
    default_params();
    g.init = true;
    return GF_OK;
} GF_CATCH_STATUS

int gf_finalize(void)
try {
    if (!g.init) return GF_OK;
    hipStreamSynchronize(g.stream);
    if (g.batch_pool) gf_pool_destroy(g.batch_pool);
    g.batch_pool = nullptr;
    if (g.h_stage) hipHostFree(g.h_stage);
    g.h_stage = nullptr;
    g.h_stage_bytes = 0;
    for (auto e : g.ev_step) hipEventDestroy(e);
    for (auto e : g.ev_tune) hipEventDestroy(e);
    g.ev_step.clear();
    hipEventDestroy(g.ev_begin);
    hipEventDestroy(g.ev_end);
    hipEventDestroy(g.ev_half[0]);
    hipEventDestroy(g.ev_half[1]);
    g.ev_half[0] = g.ev_half[1] = nullptr;
    delete g.workers;
    g.workers = nullptr;
    hipFree(g.d_max);
    hipHostFree(g.h_max);
    hipStreamDestroy(g.stream);
    g = Lib();
    return GF_OK;
} GF_CATCH_STATUS

int gf_check_parameters(int mwaves)
try {
    if (mwaves != 3) return fail(GF_EUNSUPPORTED, "mwaves = %d; the shallow-water solver has 3 waves", mwaves);
    return GF_OK;
} GF_CATCH_STATUS

int gf_set_params(const gf_params* p)
try {
    if (int r = need_init()) return r;
    if (!p) return fail(GF_EINVAL, "params is NULL");
    if (!(p->grav > 0.0)) return fail(GF_EINVAL, "grav must be > 0");
    if (!(p->dry_tolerance >= 0.0)) return fail(GF_EINVAL, "dry_tolerance must be >= 0");
    if (p->order != 1 && p->order != 2) return fail(GF_EINVAL, "order must be 1 or 2");
    if (p->transverse < 0 || p->transverse > 2) return fail(GF_EINVAL, "transverse must be 0, 1 or 2");
    if (p->mcapa != 0) return fail(GF_EUNSUPPORTED, "mcapa = %d; only mcapa = 0 is supported", p->mcapa);
    for (int k = 0; k < 3; k++)
        if (p->mthlim[k] < 0 || p->mthlim[k] > 5) return fail(GF_EINVAL, "mthlim[%d] = %d out of [0,5]", k, p->mthlim[k]);
    g.user = *p;
    g.P.grav = p->grav;
    g.P.dry_tolerance = p->dry_tolerance;
    g.P.manning_n = p->manning_n;
    g.P.friction_depth = p->friction_depth;
    g.P.friction_forcing = p->friction_forcing;
    g.P.src_term = p->src_term;
    g.P.order = p->order;
    g.P.transverse = p->transverse;
    for (int k = 0; k < 3; k++) g.P.mthlim[k] = p->mthlim[k];
    derive_params(g.P);
    g.params_gen += 1;
    return GF_OK;
} GF_CATCH_STATUS

int gf_get_params(gf_params* p)
try {
    if (int r = need_init()) return r;
    if (!p) return fail(GF_EINVAL, "params is NULL");
    *p = g.user;
    return GF_OK;
} GF_CATCH_STATUS

const char* gf_last_error(void) { return g_err.c_str(); }

long long gf_pool_bytes(int mx, int my, int mbc, int capacity)
{
    const long long np2 = (long long)(mx + 2 * mbc) * (my + 2 * mbc);
    const long long per = (2 * 3 * np2 + np2) * 8 + (long long)sizeof(PatchDesc) + 8 /*cfl*/ + 16 /*hmax*/ +
                          8 /*cflbits*/ + 8 /*lists*/ + 8 /*legacy hmax*/ + 4 /*dry age*/ + sizeof(ScenCtrl);
    const long long staging = (long long)kChunk * 4 * np2 * 8;
    return per * capacity + staging;
}

// Forget the steady-dry history of slots [first, first + n) after a write
// of their state, descriptors or parameters from outside the batched step.
static int reset_dry_age(gf_pool* p, int first, int n)
{
    if (n > 0) HIPCHK(hipMemsetAsync(p->dry_age + first, 0, (size_t)n * sizeof(int), g.stream));
    return GF_OK;
}

static void free_hydrographs(gf_pool* p)
{
    for (double*& m : p->hyd_mem) {
        if (m) hipFree(m);
        m = nullptr;
    }
}

// Rebuild the device list of patches with a hydrograph side, and check that
// every scenario with such patches has a table and uses one side only.
static int rebuild_inflow(gf_pool* p)
{
    std::vector<int> list;
    p->inflow_error.clear();
    for (int s = 0; s < p->nscen; s++) {   // scenario ids in use: [0, nscen)
        const ScenCtrl& c = p->h_ctrl[s];
        if (c.npatch == 0) continue;
        int sides = 0;
        for (int k = c.first_patch; k < c.first_patch + c.npatch; k++) {
            bool in = false;
            for (int e = 0; e < 4; e++)
                if (p->h_desc[k].nbr[e] == GF_NB_INFLOW) {
                    sides |= 1 << e;
                    in = true;
                }
            if (in) list.push_back(k);
        }
        if (sides && (sides & (sides - 1)) && p->inflow_error.empty())
            p->inflow_error = "scenario " + std::to_string(s) + " has hydrograph inflow on more than one side";
        if (sides && c.hyd_rows == 0 && p->inflow_error.empty())
            p->inflow_error = "scenario " + std::to_string(s) + " has inflow sides but no hydrograph";
    }
    const int n = (int)list.size();
    if ((n > 0) != p->has_inflow) {
        p->has_inflow = n > 0;
        p->graph_dirty = true;
    }
    if (p->amr) {
        // level by level: the ghost update fills a level's hydrograph sides
        // before interpolating from it to the next finer level
        std::stable_sort(list.begin(), list.end(), [&](int a, int b) { return p->h_level[a] < p->h_level[b]; });
        std::vector<int> rng(2 * kMaxLevels);
        unsigned levels = 0;
        for (int l = 0, k = 0; l < kMaxLevels; l++) {
            rng[2 * l] = k;
            while (k < n && p->h_level[list[k]] == l) k++;
            rng[2 * l + 1] = k;
            if (k > rng[2 * l]) levels |= 1u << l;
        }
        if (levels != p->inflow_levels) {
            p->inflow_levels = levels;
            p->graph_dirty = true;
        }
        HIPCHK(hipMemcpyAsync(p->d_ranges + range_inflow(0), rng.data(), sizeof(int) * rng.size(),
                              hipMemcpyHostToDevice, g.stream));
    }
    if (n) HIPCHK(hipMemcpyAsync(p->inflow_list, list.data(), sizeof(int) * n, hipMemcpyHostToDevice, g.stream));
    HIPCHK(hipMemcpyAsync(p->inflow_n, &n, sizeof(int), hipMemcpyHostToDevice, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    return GF_OK;
}

gf_pool* gf_pool_create(int mx, int my, int mbc, int capacity)
try {
    if (need_init()) return nullptr;
    if (mbc != gf::MBC) {
        fail(GF_EUNSUPPORTED, "mbc = %d; only mbc = 2 is supported", mbc);
        return nullptr;
    }
    if (!gf::supported_shape(mx, my)) {
        fail(GF_EUNSUPPORTED, "patch shape %dx%d not supported (8x8, 16x16, 32x32, 64x64)", mx, my);
        return nullptr;
    }
    if (capacity <= 0) {
        fail(GF_EINVAL, "capacity must be > 0");
        return nullptr;
    }
    // freed by gf_pool_destroy if anything below fails or throws
    std::unique_ptr<gf_pool, int (*)(gf_pool*)> hold(new gf_pool(), gf_pool_destroy);
    gf_pool* p = hold.get();
    p->mx = mx;
    p->my = my;
    p->mbc = mbc;
    p->capacity = capacity;
    p->np2 = (size_t)(mx + 2 * mbc) * (my + 2 * mbc);
    const size_t qbytes = (size_t)capacity * 3 * p->np2 * sizeof(double);
    const size_t chunk = std::min<size_t>(kChunk, capacity);
    bool ok = hipMalloc(&p->q[0], qbytes) == hipSuccess && hipMalloc(&p->q[1], qbytes) == hipSuccess &&
              hipMalloc(&p->aux, (size_t)capacity * p->np2 * sizeof(double)) == hipSuccess &&
              hipMalloc(&p->desc, (size_t)capacity * sizeof(PatchDesc)) == hipSuccess &&
              hipMalloc(&p->cfl, (size_t)capacity * sizeof(double)) == hipSuccess &&
              hipMalloc(&p->hmax, (size_t)capacity * sizeof(double)) == hipSuccess &&
              hipMalloc(&p->ctrl, (size_t)capacity * sizeof(ScenCtrl)) == hipSuccess &&
              hipMalloc(&p->d_stage, chunk * 4 * p->np2 * sizeof(double)) == hipSuccess &&
              hipHostMalloc(&p->h_stage, chunk * 4 * p->np2 * sizeof(double), hipHostMallocDefault) == hipSuccess &&
              hipHostMalloc(&p->h_cfl, (size_t)capacity * sizeof(double), hipHostMallocDefault) == hipSuccess &&
              hipMalloc(&p->hmax2, (size_t)2 * capacity * sizeof(double)) == hipSuccess &&
              hipMalloc(&p->cflbits, (size_t)capacity * sizeof(unsigned long long)) == hipSuccess &&
              hipMalloc(&p->list_wet, (size_t)capacity * sizeof(int)) == hipSuccess &&
              hipMalloc(&p->list_dry, (size_t)capacity * sizeof(int)) == hipSuccess &&
              hipMalloc(&p->counters, gf::NCOUNTERS * sizeof(int)) == hipSuccess &&
              hipMalloc(&p->stats, 4 * sizeof(long long)) == hipSuccess &&
              hipMalloc(&p->d_args, sizeof(gf::BatchArgs)) == hipSuccess &&
              hipMalloc(&p->tile_cfl, (size_t)capacity * sizeof(unsigned long long)) == hipSuccess &&
              hipMalloc(&p->tile_h, (size_t)capacity * sizeof(unsigned long long)) == hipSuccess &&
              hipMalloc(&p->tile_cnt, (size_t)capacity * sizeof(int)) == hipSuccess &&
              hipMalloc(&p->dry_age, (size_t)capacity * sizeof(int)) == hipSuccess &&
              hipMalloc(&p->inflow_list, (size_t)capacity * sizeof(int)) == hipSuccess &&
              hipMalloc(&p->inflow_n, sizeof(int)) == hipSuccess;
    if (ok) p->stage_doubles = chunk * 4 * p->np2;
    if (ok)
        ok = hipMemsetAsync(p->q[0], 0, qbytes, g.stream) == hipSuccess &&
             hipMemsetAsync(p->q[1], 0, qbytes, g.stream) == hipSuccess &&
             hipMemsetAsync(p->aux, 0, (size_t)capacity * p->np2 * sizeof(double), g.stream) == hipSuccess &&
             hipMemsetAsync(p->desc, 0, (size_t)capacity * sizeof(PatchDesc), g.stream) == hipSuccess &&
             hipMemsetAsync(p->cfl, 0, (size_t)capacity * sizeof(double), g.stream) == hipSuccess &&
             hipMemsetAsync(p->hmax, 0, (size_t)capacity * sizeof(double), g.stream) == hipSuccess &&
             hipMemsetAsync(p->ctrl, 0, (size_t)capacity * sizeof(ScenCtrl), g.stream) == hipSuccess &&
             hipMemsetAsync(p->hmax2, 0, (size_t)2 * capacity * sizeof(double), g.stream) == hipSuccess &&
             hipMemsetAsync(p->cflbits, 0, (size_t)capacity * sizeof(unsigned long long), g.stream) == hipSuccess &&
             hipMemsetAsync(p->counters, 0, gf::NCOUNTERS * sizeof(int), g.stream) == hipSuccess &&
             hipMemsetAsync(p->stats, 0, 4 * sizeof(long long), g.stream) == hipSuccess &&
             hipMemsetAsync(p->tile_cfl, 0, (size_t)capacity * sizeof(unsigned long long), g.stream) == hipSuccess &&
             hipMemsetAsync(p->tile_h, 0, (size_t)capacity * sizeof(unsigned long long), g.stream) == hipSuccess &&
             hipMemsetAsync(p->tile_cnt, 0, (size_t)capacity * sizeof(int), g.stream) == hipSuccess &&
             hipMemsetAsync(p->dry_age, 0, (size_t)capacity * sizeof(int), g.stream) == hipSuccess &&
             hipMemsetAsync(p->inflow_list, 0, (size_t)capacity * sizeof(int), g.stream) == hipSuccess &&
             hipMemsetAsync(p->inflow_n, 0, sizeof(int), g.stream) == hipSuccess &&
             hipStreamSynchronize(g.stream) == hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        fail(GF_ENOMEM, "pool allocation of %d patches (%lld bytes) failed", capacity,
             gf_pool_bytes(mx, my, mbc, capacity));
        return nullptr;
    }
    p->h_desc.resize(capacity);
    std::memset(p->h_desc.data(), 0, sizeof(PatchDesc) * capacity);
    for (auto& d : p->h_desc) {
        d.scen = -1;
        for (int k = 0; k < 8; k++) d.nbr[k] = GF_NB_STORED;
    }
    if (hipMemcpyAsync(p->desc, p->h_desc.data(), sizeof(PatchDesc) * capacity, hipMemcpyHostToDevice,
                       g.stream) != hipSuccess ||
        hipStreamSynchronize(g.stream) != hipSuccess) {
        fail(GF_EHIP, "descriptor upload failed");
        return nullptr;
    }
    p->h_ctrl.assign(capacity, ScenCtrl{});
    p->hyd_mem.assign(capacity, nullptr);
    p->h_level.assign(capacity, 0);
    return hold.release();
} GF_CATCH_NULL

int gf_pool_destroy(gf_pool* p)
try {
    if (!p) return GF_OK;
    if (g.stream) hipStreamSynchronize(g.stream);
    if (p->graph_exec) hipGraphExecDestroy(p->graph_exec);
    if (p->graph) hipGraphDestroy(p->graph);
    hipFree(p->q[0]);
    hipFree(p->q[1]);
    hipFree(p->aux);
    hipFree(p->desc);
    hipFree(p->cfl);
    hipFree(p->hmax);
    hipFree(p->ctrl);
    hipFree(p->d_stage);
    hipFree(p->hmax2);
    hipFree(p->cflbits);
    hipFree(p->list_wet);
    hipFree(p->list_dry);
    hipFree(p->counters);
    hipFree(p->stats);
    hipFree(p->d_args);
    hipFree(p->tile_cfl);
    hipFree(p->tile_h);
    hipFree(p->tile_cnt);
    hipFree(p->dry_age);
    hipFree(p->tag_buf);
    hipFree(p->inflow_list);
    hipFree(p->inflow_n);
    hipFree(p->d_links);
    hipFree(p->d_ranges);
    for (double* m : p->hyd_mem) hipFree(m);
    if (p->h_stage) hipHostFree(p->h_stage);
    if (p->h_cfl) hipHostFree(p->h_cfl);
    delete p;
    return GF_OK;
} GF_CATCH_STATUS

int gf_pool_set_patches(gf_pool* p, int n, const double* dx, const double* dy, const double* xlo,
                        const double* ylo, const int* scen, const int* nbr)
try {
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    if (n < 0 || n > p->capacity) return fail(GF_EINVAL, "n = %d outside [0, %d]", n, p->capacity);
    if (n > 0 && (!dx || !dy)) return fail(GF_EINVAL, "dx/dy are NULL");
    for (int k = 0; k < n; k++) {
        PatchDesc& d = p->h_desc[k];
        if (!(dx[k] > 0.0) || !(dy[k] > 0.0)) return fail(GF_EINVAL, "patch %d: dx, dy must be > 0", k);
        d.dx = dx[k];
        d.dy = dy[k];
        d.xlo = xlo ? xlo[k] : 0.0;
        d.ylo = ylo ? ylo[k] : 0.0;
        if (scen && scen[k] != -1)
            return fail(GF_EINVAL, "patch %d: scenario ids are assigned by gf_pool_add_scenario (pass -1)", k);
        d.scen = -1;
        for (int s = 0; s < 8; s++) {
            const int v = nbr ? nbr[8 * k + s] : GF_NB_STORED;
            if (v >= n || v < GF_NB_INFLOW || (s >= 4 && v == GF_NB_INFLOW))
                return fail(GF_EINVAL, "patch %d: neighbour entry %d = %d invalid", k, s, v);
            d.nbr[s] = v;
        }
        d.pad = 0;
    }
    for (int k = n; k < p->capacity; k++) p->h_desc[k].scen = -1;
    p->n = n;
    p->nscen = 0;
    p->h_ctrl.assign(p->capacity, ScenCtrl{});
    free_hydrographs(p);
    HIPCHK(hipMemcpyAsync(p->desc, p->h_desc.data(), sizeof(PatchDesc) * p->capacity, hipMemcpyHostToDevice, g.stream));
    HIPCHK(hipMemsetAsync(p->ctrl, 0, sizeof(ScenCtrl) * p->capacity, g.stream));
    if (int r = reset_dry_age(p, 0, p->capacity)) return r;
    HIPCHK(hipStreamSynchronize(g.stream));
    p->graph_dirty = true;
    p->args_dirty = true;
    return rebuild_inflow(p);
} GF_CATCH_STATUS

int gf_pool_describe(gf_pool* p, int first, int n, const double* dx, const double* dy, const double* xlo,
                     const double* ylo, const int* nbr)
try {
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    if (first < 0 || n <= 0 || first + n > p->capacity) return fail(GF_EINVAL, "slots [%d,%d) outside the pool", first, first + n);
    if (!dx || !dy || !nbr) return fail(GF_EINVAL, "NULL argument");
    for (int k = 0; k < n; k++) {
        PatchDesc& d = p->h_desc[first + k];
        if (!(dx[k] > 0.0) || !(dy[k] > 0.0)) return fail(GF_EINVAL, "patch %d: dx, dy must be > 0", first + k);
        for (int s = 0; s < 8; s++) {
            const int v = nbr[8 * k + s];
            if (v >= p->capacity || v < GF_NB_INFLOW || (s >= 4 && v == GF_NB_INFLOW))
                return fail(GF_EINVAL, "patch %d: neighbour entry %d = %d invalid", first + k, s, v);
        }
        if (d.scen >= 0) return fail(GF_EINVAL, "patch %d belongs to scenario %d", first + k, d.scen);
    }
    for (int k = 0; k < n; k++) {
        PatchDesc& d = p->h_desc[first + k];
        d.dx = dx[k];
        d.dy = dy[k];
        d.xlo = xlo ? xlo[k] : 0.0;
        d.ylo = ylo ? ylo[k] : 0.0;
        d.scen = -1;
        for (int s = 0; s < 8; s++) d.nbr[s] = nbr[8 * k + s];
    }
    if (first + n > p->n) p->n = first + n;
    HIPCHK(hipMemcpyAsync(p->desc + first, p->h_desc.data() + first, sizeof(PatchDesc) * n, hipMemcpyHostToDevice, g.stream));
    if (int r = reset_dry_age(p, first, n)) return r;
    HIPCHK(hipStreamSynchronize(g.stream));
    return GF_OK;
} GF_CATCH_STATUS

}  // extern "C"

// Host copies of n items spread over up to 16 host threads (the caller's and
// up to 15 kept workers).
static void host_parallel(int n, const std::function<void(int)>& f)
{
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (nt <= 1 || n <= 1) {
        for (int k = 0; k < n; k++) f(k);
        return;
    }
    if (!g.workers) {
        // no exception may cross the C ABI: without threads the copies run here
        try {
            g.workers = new HostWorkers(nt - 1);
        } catch (...) {
            g.workers = nullptr;
        }
        if (!g.workers) {
            for (int k = 0; k < n; k++) f(k);
            return;
        }
    }
    g.workers->run(n, f);
}

// Split of a pool's staging buffers (stage_doubles each, at least
// min(kChunk, capacity) * 4 * np2) into nbuf parts of `part` patches of `per`
// doubles: two parts when two patches fit, so that the DMA and (un)pack kernel
// of one part overlap the host copies into or out of the other.
struct StageSplit {
    int nbuf, part;
};
static StageSplit stage_split(const gf_pool* p, size_t per)
{
    const size_t fit = per ? p->stage_doubles / per : 0;
    return fit >= 2 ? StageSplit{2, (int)(fit / 2)} : StageSplit{1, (int)fit};
}

// Grow the staging buffers so that two patches of `per` doubles fit (a small
// pool with several aux fields per cell); they never shrink.  The new
// buffers are allocated before the old ones are freed, so a failed grow
// leaves the pool with the staging it had.
static int ensure_stage(gf_pool* p, size_t per)
{
    if (p->stage_doubles >= 2 * per) return GF_OK;
    double* d = nullptr;
    double* h = nullptr;
    if (hipMalloc(&d, 2 * per * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(GF_ENOMEM, "staging buffers of %zu bytes", 2 * per * sizeof(double));
    }
    if (hipHostMalloc(&h, 2 * per * sizeof(double), hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        hipFree(d);
        return fail(GF_ENOMEM, "staging buffers of %zu bytes", 2 * per * sizeof(double));
    }
    HIPCHK(hipStreamSynchronize(g.stream));
    HIPCHK(hipFree(p->d_stage));
    HIPCHK(hipHostFree(p->h_stage));
    p->d_stage = d;
    p->h_stage = h;
    p->stage_doubles = 2 * per;
    return GF_OK;
}

// Host patches -> pool slots [first, first+n) of buffer buf (and of buffer
// buf2 as well when buf2 >= 0) through the pinned and device staging
// buffers. put(k, q, aux) copies host patch k's 3*np2 values of q (and
// maux*np2 of aux when with_aux) into the pinned staging.
template <class Put>
static int staged_upload(gf_pool* p, int first, int n, int buf, int maux, bool with_aux, Put&& put, int buf2 = -1)
{
    const size_t np2 = p->np2;
    const size_t per = (3 + (with_aux ? (size_t)maux : 0)) * np2;
    if (int r = ensure_stage(p, per)) return r;
    const StageSplit sp = stage_split(p, per);
    if (sp.part <= 0) return fail(GF_ESTATE, "staging buffers hold no patch");
    for (int off = 0, c = 0; off < n; off += sp.part, c++) {
        const int b = c % sp.nbuf, m = std::min(sp.part, n - off);
        double* hs = p->h_stage + (size_t)b * sp.part * per;
        double* ds = p->d_stage + (size_t)b * sp.part * per;
        if (c >= sp.nbuf) HIPCHK(hipEventSynchronize(g.ev_half[b]));
        const size_t qd = (size_t)m * 3 * np2;
        host_parallel(m, [&](int k) {
            put(off + k, hs + (size_t)k * 3 * np2, with_aux ? hs + qd + (size_t)k * maux * np2 : nullptr);
        });
        HIPCHK(hipMemcpyAsync(ds, hs, (qd + (with_aux ? (size_t)m * maux * np2 : 0)) * sizeof(double),
                              hipMemcpyHostToDevice, g.stream));
        HIPCHK(gf::launch_pack(ds, with_aux ? ds + qd : nullptr, p->q[buf], buf2 >= 0 ? p->q[buf2] : nullptr,
                               p->aux, first + off, m, (int)np2, maux, g.stream));
        HIPCHK(hipEventRecord(g.ev_half[b], g.stream));
    }
    return GF_OK;
}

// Pool slots [first, first+n) of buffer buf -> host, the reverse pipeline:
// get(k, q) copies patch k's 3*np2 values out of the pinned staging.
template <class Get>
static int staged_download(gf_pool* p, int first, int n, int buf, Get&& get)
{
    const size_t np2 = p->np2;
    const size_t per = 3 * np2;
    const StageSplit sp = stage_split(p, per);
    if (sp.part <= 0) return fail(GF_ESTATE, "staging buffers hold no patch");
    const int nch = (n + sp.part - 1) / sp.part;
    auto enqueue = [&](int c) -> int {
        const int b = c % sp.nbuf, off = c * sp.part, m = std::min(sp.part, n - off);
        double* hs = p->h_stage + (size_t)b * sp.part * per;
        double* ds = p->d_stage + (size_t)b * sp.part * per;
        HIPCHK(gf::launch_unpack(ds, p->q[buf], first + off, m, (int)np2, g.stream));
        HIPCHK(hipMemcpyAsync(hs, ds, (size_t)m * per * sizeof(double), hipMemcpyDeviceToHost, g.stream));
        HIPCHK(hipEventRecord(g.ev_half[b], g.stream));
        return GF_OK;
    };
    for (int c = 0; c < nch; c++) {
        // with two parts, chunk c+1's unpack and DMA run while the host copies
        // chunk c out; that part's previous chunk was copied out before
        if (c == 0 || sp.nbuf == 1)
            if (int r = enqueue(c)) return r;
        if (sp.nbuf == 2 && c + 1 < nch)
            if (int r = enqueue(c + 1)) return r;
        const int b = c % sp.nbuf, off = c * sp.part, m = std::min(sp.part, n - off);
        const double* hs = p->h_stage + (size_t)b * sp.part * per;
        HIPCHK(hipEventSynchronize(g.ev_half[b]));
        host_parallel(m, [&](int k) { get(off + k, hs + (size_t)k * per); });
    }
    return GF_OK;
}

extern "C" {

int gf_pool_upload(gf_pool* p, int first, int n, int buf, const double* q_aos, const double* aux_aos, int maux)
try {
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    if (first < 0 || n < 0 || first + n > p->capacity) return fail(GF_EINVAL, "slots [%d,%d) outside the pool", first, first + n);
    if (buf != 0 && buf != 1) return fail(GF_EINVAL, "buffer must be 0 or 1");
    if (n > 0 && !q_aos) return fail(GF_EINVAL, "q is NULL");
    if (aux_aos && maux < 1) return fail(GF_EINVAL, "maux must be >= 1");
    const size_t np2 = p->np2;
    if (int r = staged_upload(p, first, n, buf, aux_aos ? maux : 1, aux_aos != nullptr,
                              [&](int k, double* q, double* a) {
                                  std::memcpy(q, q_aos + (size_t)k * 3 * np2, 3 * np2 * sizeof(double));
                                  if (a) std::memcpy(a, aux_aos + (size_t)k * maux * np2, maux * np2 * sizeof(double));
                              }))
        return r;
    // interior depth maxima of the uploaded state (classification of the first attempt)
    HIPCHK(gf::launch_hmax(p->mx, p->my, p->q[buf], p->hmax2 + (size_t)buf * p->capacity, first, n, g.stream));
    if (int r = reset_dry_age(p, first, n)) return r;
    HIPCHK(hipStreamSynchronize(g.stream));
    return GF_OK;
} GF_CATCH_STATUS

int gf_pool_download(gf_pool* p, int first, int n, int buf, double* q_aos)
try {
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    if (first < 0 || n < 0 || first + n > p->capacity) return fail(GF_EINVAL, "slots [%d,%d) outside the pool", first, first + n);
    if (buf != 0 && buf != 1) return fail(GF_EINVAL, "buffer must be 0 or 1");
    if (n > 0 && !q_aos) return fail(GF_EINVAL, "q is NULL");
    const size_t np2 = p->np2;
    return staged_download(p, first, n, buf, [&](int k, const double* q) {
        std::memcpy(q_aos + (size_t)k * 3 * np2, q, 3 * np2 * sizeof(double));
    });
} GF_CATCH_STATUS

int gf_pool_download_aux(gf_pool* p, int first, int n, double* aux_aos)
try {
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    if (first < 0 || n < 0 || first + n > p->capacity) return fail(GF_EINVAL, "slots [%d,%d) outside the pool", first, first + n);
    if (n > 0 && !aux_aos) return fail(GF_EINVAL, "aux is NULL");
    if (n == 0) return GF_OK;
    HIPCHK(hipMemcpyAsync(aux_aos, p->aux + (size_t)first * p->np2, sizeof(double) * (size_t)n * p->np2,
                          hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    return GF_OK;
} GF_CATCH_STATUS

static gf::InflowArgs inflow_args(gf_pool* p, int first, int last, int buf, double t)
{
    gf::InflowArgs a{};
    a.desc = p->desc;
    a.ctrl = p->ctrl;
    a.q0 = p->q[0];
    a.q1 = p->q[1];
    a.aux = p->aux;
    a.list = p->inflow_list;
    a.list_n = p->inflow_n;
    a.first = first;
    a.last = last;
    a.buf = buf;
    a.t = t;
    a.P = g.P;
    return a;
}

// Workgroups of k_inflow (grid-stride over the inflow list): a few per CU.
static int inflow_grid() { return 4 * (g.num_cu > 0 ? g.num_cu : 64); }

int gf_pool_ghost_fill(gf_pool* p, int first, int n, int buf, double t)
try {
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    if (first < 0 || n < 0 || first + n > p->n) return fail(GF_EINVAL, "patches [%d,%d) outside [0,%d)", first, first + n, p->n);
    if (buf != 0 && buf != 1) return fail(GF_EINVAL, "buffer must be 0 or 1");
    if (!p->inflow_error.empty()) return fail(GF_EINVAL, "%s", p->inflow_error.c_str());
    gf::GhostArgs a{p->desc, first, n, p->q[0], p->q[1], nullptr, buf};
    HIPCHK(gf::launch_ghost(p->mx, p->my, a, g.stream));
    HIPCHK(gf::launch_inflow(p->mx, p->my, inflow_args(p, first, first + n, buf, t), inflow_grid(), g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    return GF_OK;
} GF_CATCH_STATUS

static int pool_step_impl(gf_pool* p, int first, int n, int src, int dst, double dt, double* cfl,
                          double* maxcfl, bool plain)
{
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    if (first < 0 || n < 0 || first + n > p->n) return fail(GF_EINVAL, "patches [%d,%d) outside [0,%d)", first, first + n, p->n);
    if ((src != 0 && src != 1) || (dst != 0 && dst != 1)) return fail(GF_EINVAL, "buffers must be 0 or 1");
    if (!(dt > 0.0)) return fail(GF_EINVAL, "dt must be > 0");
    if (src == dst) {
        // A tiled variant steps a patch with several workgroups, each reading
        // its halo from the neighbouring tiles' cells, which the others
        // overwrite in place: only one workgroup per patch may step in place.
        gf::StepVariant v;
        if (!gf::get_variant(p->mx, p->my, p->variant, &v)) return fail(GF_ESTATE, "no step variant %d", p->variant);
        if (v.tx != p->mx || v.ty != p->my)
            return fail(GF_EUNSUPPORTED,
                        "an in-place step (src == dst) needs a step variant with one tile per patch; variant %d "
                        "splits %dx%d patches into %dx%d tiles (step into the other buffer instead)",
                        p->variant, p->mx, p->my, v.tx, v.ty);
    }
    double* q0 = (src == dst) ? p->q[src] : p->q[0];
    double* q1 = (src == dst) ? nullptr : p->q[1];
    const int cur = (src == dst) ? 0 : src;
    const int launches = plain ? n : (n > 0 ? 1 : 0);
    if (int r = ensure_step_events(std::max(1, launches))) return r;
    if (int r = reset_dry_age(p, first, n)) return r;
    HIPCHK(hipEventRecord(g.ev_begin, g.stream));
    if (plain) {
        for (int k = 0; k < n; k++) {
            HIPCHK(hipEventRecord(g.ev_step[2 * k], g.stream));
            HIPCHK(gf::launch_step(p->mx, p->my, p->variant, step_args(p, first + k, 1, q0, q1, nullptr, dt, cur), g.stream));
            HIPCHK(hipEventRecord(g.ev_step[2 * k + 1], g.stream));
        }
    } else if (n > 0) {
        HIPCHK(hipEventRecord(g.ev_step[0], g.stream));
        HIPCHK(gf::launch_step(p->mx, p->my, p->variant, step_args(p, first, n, q0, q1, nullptr, dt, cur), g.stream));
        HIPCHK(hipEventRecord(g.ev_step[1], g.stream));
    }
    if (n > 0) HIPCHK(hipMemcpyAsync(p->h_cfl, p->cfl + first, sizeof(double) * n, hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipEventRecord(g.ev_end, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, g.ev_begin, g.ev_end));
    g.last_total_ms = ms;
    g.last_step_ms = 0.0;
    for (int k = 0; k < launches; k++) {
        HIPCHK(hipEventElapsedTime(&ms, g.ev_step[2 * k], g.ev_step[2 * k + 1]));
        g.last_step_ms += ms;
    }
    g.last_launches = launches;
    double m = 0.0;
    for (int k = 0; k < n; k++) {
        if (cfl) cfl[k] = p->h_cfl[k];
        m = max_nan(m, p->h_cfl[k]);
    }
    if (maxcfl) *maxcfl = m;
    return GF_OK;
}

int gf_pool_step(gf_pool* p, int first, int n, int src, int dst, double dt, double* cfl, double* maxcfl)
try {
    return pool_step_impl(p, first, n, src, dst, dt, cfl, maxcfl, false);
} GF_CATCH_STATUS

int gf_pool_step_plain(gf_pool* p, int first, int n, int src, int dst, double dt, double* cfl, double* maxcfl)
try {
    return pool_step_impl(p, first, n, src, dst, dt, cfl, maxcfl, true);
} GF_CATCH_STATUS

int gf_pool_add_scenario(gf_pool* p, int first_patch, int npatch, double t0, double dt0,
                         double desired_cfl, double max_cfl, int target_steps, double t_end)
try {
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    if (first_patch < 0 || npatch <= 0 || first_patch + npatch > p->n)
        return fail(GF_EINVAL, "scenario patches [%d,%d) outside the described slots [0,%d)", first_patch,
                    first_patch + npatch, p->n);
    if (!(dt0 > 0.0) || !(desired_cfl > 0.0) || !(max_cfl > 0.0)) return fail(GF_EINVAL, "dt0, desired_cfl, max_cfl must be > 0");
    for (int k = first_patch; k < first_patch + npatch; k++)
        if (p->h_desc[k].scen >= 0) return fail(GF_EINVAL, "patch %d already belongs to scenario %d", k, p->h_desc[k].scen);
    int id = -1;
    for (int s = 0; s < p->capacity; s++)
        if (p->h_ctrl[s].npatch == 0) {
            id = s;
            break;
        }
    if (id < 0) return fail(GF_EINVAL, "too many scenarios");
    ScenCtrl c{};
    c.t = t0;
    c.dt = dt0;
    c.maxcfl = 0.0;
    c.desired_cfl = desired_cfl;
    c.max_cfl = max_cfl;
    c.t_end = t_end;
    c.last_dt = 0.0;
    c.manning_n = g.P.manning_n;
    c.cur = 0;
    c.nsteps = 0;
    c.target_steps = target_steps;
    c.retakes = 0;
    c.active = 1;
    c.first_patch = first_patch;
    c.npatch = npatch;
    c.attempts = 0;
    p->h_ctrl[id] = c;
    for (int k = first_patch; k < first_patch + npatch; k++) p->h_desc[k].scen = id;
    HIPCHK(hipMemcpyAsync(p->desc + first_patch, p->h_desc.data() + first_patch, sizeof(PatchDesc) * npatch,
                          hipMemcpyHostToDevice, g.stream));
    HIPCHK(hipMemcpyAsync(p->ctrl + id, &p->h_ctrl[id], sizeof(ScenCtrl), hipMemcpyHostToDevice, g.stream));
    if (int r = reset_dry_age(p, first_patch, npatch)) return r;
    HIPCHK(hipStreamSynchronize(g.stream));
    if (id + 1 > p->nscen) p->nscen = id + 1;
    if (int r = rebuild_inflow(p)) return r;
    return id;
} GF_CATCH_STATUS

int gf_pool_remove_scenario(gf_pool* p, int s)
try {
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    if (s < 0 || s >= p->capacity || p->h_ctrl[s].npatch == 0) return fail(GF_EINVAL, "scenario %d unknown", s);
    const int first = p->h_ctrl[s].first_patch, n = p->h_ctrl[s].npatch;
    for (int k = first; k < first + n; k++) p->h_desc[k].scen = -1;
    p->h_ctrl[s] = ScenCtrl{};
    HIPCHK(hipMemcpyAsync(p->desc + first, p->h_desc.data() + first, sizeof(PatchDesc) * n, hipMemcpyHostToDevice, g.stream));
    HIPCHK(hipMemcpyAsync(p->ctrl + s, &p->h_ctrl[s], sizeof(ScenCtrl), hipMemcpyHostToDevice, g.stream));
    if (int r = reset_dry_age(p, first, n)) return r;
    HIPCHK(hipStreamSynchronize(g.stream));
    if (p->hyd_mem[s]) HIPCHK(hipFree(p->hyd_mem[s]));
    p->hyd_mem[s] = nullptr;
    return rebuild_inflow(p);
} GF_CATCH_STATUS

int gf_pool_set_scenario_inflow(gf_pool* p, int s, int nrows, const double* time, const double* hu,
                                double b, double x0, double y0)
try {
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    if (s < 0 || s >= p->capacity || p->h_ctrl[s].npatch == 0) return fail(GF_EINVAL, "scenario %d unknown", s);
    if (nrows < 1 || !time || !hu) return fail(GF_EINVAL, "hydrograph needs >= 1 rows of time and discharge");
    if (!(b >= 0.0)) return fail(GF_EINVAL, "channel width must be >= 0");
    double* d = nullptr;
    HIPCHK(hipMalloc(&d, sizeof(double) * 2 * (size_t)nrows));
    if (hipMemcpy(d, time, sizeof(double) * nrows, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d + nrows, hu, sizeof(double) * nrows, hipMemcpyHostToDevice) != hipSuccess) {
        hipFree(d);
        return fail(GF_EHIP, "hydrograph upload failed");
    }
    HIPCHK(hipMemcpyAsync(&p->h_ctrl[s], p->ctrl + s, sizeof(ScenCtrl), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    ScenCtrl& c = p->h_ctrl[s];
    c.hyd = d;
    c.hyd_rows = nrows;
    c.hyd_b = b;
    c.hyd_x0 = x0;
    c.hyd_y0 = y0;
    HIPCHK(hipMemcpyAsync(p->ctrl + s, &c, sizeof(ScenCtrl), hipMemcpyHostToDevice, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    if (p->hyd_mem[s]) HIPCHK(hipFree(p->hyd_mem[s]));
    p->hyd_mem[s] = d;
    return rebuild_inflow(p);
} GF_CATCH_STATUS

int gf_pool_set_scenario_friction(gf_pool* p, int s, double manning_n)
try {
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    if (s < 0 || s >= p->capacity || p->h_ctrl[s].npatch == 0) return fail(GF_EINVAL, "scenario %d unknown", s);
    if (!(manning_n >= 0.0)) return fail(GF_EINVAL, "manning_n must be >= 0");
    HIPCHK(hipMemcpyAsync(&p->h_ctrl[s], p->ctrl + s, sizeof(ScenCtrl), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    p->h_ctrl[s].manning_n = manning_n;
    HIPCHK(hipMemcpyAsync(p->ctrl + s, &p->h_ctrl[s], sizeof(ScenCtrl), hipMemcpyHostToDevice, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    return GF_OK;
} GF_CATCH_STATUS

int gf_pool_set_scenario_stop(gf_pool* p, int s, int target_steps, double t_end)
try {
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    if (s < 0 || s >= p->capacity || p->h_ctrl[s].npatch == 0) return fail(GF_EINVAL, "scenario %d unknown", s);
    HIPCHK(hipMemcpyAsync(&p->h_ctrl[s], p->ctrl + s, sizeof(ScenCtrl), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    ScenCtrl& c = p->h_ctrl[s];
    c.target_steps = target_steps;
    c.t_end = t_end;
    const bool done = (target_steps > 0 && c.nsteps >= target_steps) || (t_end > 0.0 && c.t >= t_end);
    c.active = (c.status == gf::SCEN_OK && !done) ? 1 : 0;
    HIPCHK(hipMemcpyAsync(p->ctrl + s, &c, sizeof(ScenCtrl), hipMemcpyHostToDevice, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    return GF_OK;
} GF_CATCH_STATUS

int gf_pool_set_scenario_dt_max(gf_pool* p, int s, double dt_max)
try {
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    if (s < 0 || s >= p->capacity || p->h_ctrl[s].npatch == 0) return fail(GF_EINVAL, "scenario %d unknown", s);
    if (!(dt_max >= 0.0) || dt_max - dt_max != 0.0) return fail(GF_EINVAL, "dt_max must be finite and >= 0");
    HIPCHK(hipMemcpyAsync(&p->h_ctrl[s], p->ctrl + s, sizeof(ScenCtrl), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    p->h_ctrl[s].dt_max = dt_max;
    HIPCHK(hipMemcpyAsync(p->ctrl + s, &p->h_ctrl[s], sizeof(ScenCtrl), hipMemcpyHostToDevice, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    return GF_OK;
} GF_CATCH_STATUS

int gf_pool_clear_scenarios(gf_pool* p)
try {
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    for (int s = 0; s < p->capacity; s++)
        if (p->h_ctrl[s].npatch > 0)
            for (int k = p->h_ctrl[s].first_patch; k < p->h_ctrl[s].first_patch + p->h_ctrl[s].npatch; k++)
                p->h_desc[k].scen = -1;
    p->nscen = 0;
    p->h_ctrl.assign(p->capacity, ScenCtrl{});
    HIPCHK(hipMemcpyAsync(p->desc, p->h_desc.data(), sizeof(PatchDesc) * p->capacity, hipMemcpyHostToDevice, g.stream));
    HIPCHK(hipMemsetAsync(p->ctrl, 0, sizeof(ScenCtrl) * p->capacity, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    free_hydrographs(p);
    if (int r = rebuild_inflow(p)) return r;
    HIPCHK(hipMemsetAsync(p->cflbits, 0, sizeof(unsigned long long) * p->capacity, g.stream));
    if (int r = reset_dry_age(p, 0, p->capacity)) return r;
    HIPCHK(hipStreamSynchronize(g.stream));
    p->graph_dirty = true;
    p->args_dirty = true;
    return GF_OK;
} GF_CATCH_STATUS

int gf_pool_set_scenario_active(gf_pool* p, int s, int active)
try {
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    if (s < 0 || s >= p->capacity || p->h_ctrl[s].npatch == 0) return fail(GF_EINVAL, "scenario %d unknown", s);
    HIPCHK(hipMemcpyAsync(&p->h_ctrl[s], p->ctrl + s, sizeof(ScenCtrl), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    p->h_ctrl[s].active = active ? 1 : 0;
    HIPCHK(hipMemcpyAsync(p->ctrl + s, &p->h_ctrl[s], sizeof(ScenCtrl), hipMemcpyHostToDevice, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    return GF_OK;
} GF_CATCH_STATUS

int gf_pool_get_scenario(gf_pool* p, int s, gf_scenario_info* info)
try {
    if (!p || !info) return fail(GF_EINVAL, "pool or info is NULL");
    if (s < 0 || s >= p->capacity || p->h_ctrl[s].npatch == 0) return fail(GF_EINVAL, "scenario %d unknown", s);
    ScenCtrl c;
    HIPCHK(hipMemcpyAsync(&c, p->ctrl + s, sizeof(ScenCtrl), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    info->t = c.t;
    info->dt = c.dt;
    info->maxcfl = c.maxcfl;
    info->last_dt = c.last_dt;
    info->nsteps = c.nsteps;
    info->retakes = c.retakes;
    info->attempts = c.attempts;
    info->active = c.active;
    info->cur = c.cur;
    info->status = c.status;
    return GF_OK;
} GF_CATCH_STATUS

int gf_pool_get_scenarios(gf_pool* p, int first, int n, gf_scenario_info* info)
try {
    if (!p || (n > 0 && !info)) return fail(GF_EINVAL, "pool or info is NULL");
    if (first < 0 || n < 0 || first + n > p->capacity) return fail(GF_EINVAL, "scenarios [%d,%d) outside [0,%d)", first, first + n, p->capacity);
    if (n == 0) return GF_OK;
    std::vector<ScenCtrl> c(n);
    HIPCHK(hipMemcpyAsync(c.data(), p->ctrl + first, sizeof(ScenCtrl) * n, hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    for (int k = 0; k < n; k++) {
        gf_scenario_info& o = info[k];
        o.t = c[k].t;
        o.dt = c[k].dt;
        o.maxcfl = c[k].maxcfl;
        o.last_dt = c[k].last_dt;
        o.nsteps = c[k].nsteps;
        o.retakes = c[k].retakes;
        o.status = c[k].status;
        o.attempts = c[k].attempts;
        o.active = p->h_ctrl[first + k].npatch > 0 ? c[k].active : 0;
        o.cur = c[k].cur;
    }
    return GF_OK;
} GF_CATCH_STATUS

static int ensure_pair_events(std::vector<hipEvent_t>& v, size_t pairs)
{
    while (v.size() < 2 * pairs) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        v.push_back(e);
    }
    return GF_OK;
}

static gf::BatchArgs host_batch_args(gf_pool* p)
{
    gf::BatchArgs a{};
    a.desc = p->desc;
    a.n = p->n;         // the described slots; undescribed and idle ones carry scenario -1
    a.capacity = p->capacity;
    a.q0 = p->q[0];
    a.q1 = p->q[1];
    a.aux = p->aux;
    a.hmax = p->hmax2;
    a.cflbits = p->cflbits;
    a.ctrl = p->ctrl;
    a.nscen = p->nscen;  // the scenario ids in use so far; removed ones are inactive
    a.list_wet = p->list_wet;
    a.list_dry = p->list_dry;
    a.counters = p->counters;
    a.stats = p->stats;
    a.P = g.P;
    a.tile_cfl = p->tile_cfl;
    a.tile_h = p->tile_h;
    a.tile_cnt = p->tile_cnt;
    a.dry_age = p->dry_age;
    return a;
}

static int sync_batch_args(gf_pool* p)
{
    if (p->args_gen != g.params_gen) {
        // a new dry_tolerance can make a steady dry patch wet
        if (int r = reset_dry_age(p, 0, p->capacity)) return r;
        p->args_dirty = true;
        p->graph_dirty = true;
    }
    // the classify and control kernels read the slot and scenario counts
    // from the device arguments, so a captured graph stays valid when they
    // grow or shrink; only the arguments are uploaded again
    if (p->args_n != p->n || p->args_nscen != p->nscen) p->args_dirty = true;
    if (!p->args_dirty) return GF_OK;
    const gf::BatchArgs a = host_batch_args(p);
    HIPCHK(hipMemcpyAsync(p->d_args, &a, sizeof(a), hipMemcpyHostToDevice, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    p->args_dirty = false;
    p->args_gen = g.params_gen;
    p->args_n = p->n;
    p->args_nscen = p->nscen;
    return GF_OK;
}

// Workgroups of the persistent step kernel: every CU filled to the
// variant's occupancy.
static int wet_grid(gf_pool* p)
{
    if (p->wet_grid == 0) p->wet_grid = g.num_cu * gf::wet_occupancy(p->mx, p->my, p->variant);
    return p->wet_grid;
}

// The multi-level ghost update of a resident AMR pool, in ForestClaw's order
// (fclaw2d_ghost_update with time_interp off, as restated for resident
// patches): the coarse ghost cells that finer neighbours cover, averaged from
// them (all links at once: they read fine interiors only); then for each
// level from coarse to fine its hydrograph sides, and the next finer level's
// ghost cells interpolated from it.  Same-level copies and walls need no
// pass: every reader takes them from the neighbours (ghost_source).  buf < 0:
// each scenario's current buffer and time, inactive scenarios skipped (the
// form captured in the attempt graph); otherwise buffer buf at time t.
static int enqueue_ghost_update(gf_pool* p, int buf, double t)
{
    gf::RegridArgs a{};
    a.desc = p->desc;
    a.aux = p->aux;
    a.cs = p->d_links;
    a.fs = p->d_links + p->link_cap;
    a.side = p->d_links + 2 * p->link_cap;
    a.igrid = p->d_links + 3 * p->link_cap;
    a.n = kLinkGrid;
    a.dry_tolerance = g.P.dry_tolerance;
    if (buf < 0) {
        a.q = p->q[0];
        a.q1 = p->q[1];
        a.ctrl = p->ctrl;
    } else {
        a.q = p->q[buf];
    }
    if (p->nlinks > 0) {
        a.range = p->d_ranges + range_links();
        HIPCHK(gf::launch_average_ghost(p->mx, p->my, a, g.stream));
    }
    for (int l = 0; l < p->nlev; l++) {
        if (p->has_inflow && (p->inflow_levels >> l & 1u)) {
            gf::InflowArgs ia = inflow_args(p, 0, p->capacity, buf, t);
            ia.range = p->d_ranges + range_inflow(l);
            HIPCHK(gf::launch_inflow(p->mx, p->my, ia, inflow_grid(), g.stream));
        }
        if (l + 1 < p->nlev && p->nlinks > 0 && (p->interp_levels >> l & 1u)) {
            a.range = p->d_ranges + range_interp(l);
            HIPCHK(gf::launch_interpolate_ghost(p->mx, p->my, a, g.stream));
        }
    }
    return GF_OK;
}

// One attempt of every active scenario: (hydrograph ghosts, or the
// multi-level ghost update of a resident AMR pool) -> classify -> one
// persistent kernel over the active tiles and then the dry patches ->
// controller, all on one stream.  With k >= 0, timing events bracket the
// step kernel.
static int enqueue_attempt(gf_pool* p, int k)
{
    if (p->amr) {
        if (int r = enqueue_ghost_update(p, -1, 0.0)) return r;
    } else if (p->has_inflow) {
        HIPCHK(gf::launch_inflow(p->mx, p->my, inflow_args(p, 0, p->capacity, -1, 0.0), inflow_grid(), g.stream));
    }
    HIPCHK(gf::launch_classify(p->d_args, p->capacity, g.stream));
    if (k >= 0) HIPCHK(hipEventRecord(g.ev_step[2 * k], g.stream));
    HIPCHK(gf::launch_all(p->mx, p->my, p->variant, gf::default_method(g.P), p->d_args, wet_grid(p), g.stream));
    if (k >= 0) HIPCHK(hipEventRecord(g.ev_step[2 * k + 1], g.stream));
    HIPCHK(gf::launch_batch_control(p->d_args, p->capacity, g.stream));
    return GF_OK;
}

static int any_active_flag(gf_pool* p, int* any_active)
{
    if (!any_active) return GF_OK;
    if (p->nscen > 0)
        HIPCHK(hipMemcpyAsync(p->h_ctrl.data(), p->ctrl, sizeof(ScenCtrl) * p->nscen, hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    int a = 0;
    for (int s = 0; s < p->nscen; s++) a |= p->h_ctrl[s].active;
    *any_active = a;
    return GF_OK;
}

// `substeps` attempts of every active scenario (graph replays or stream
// launches), then a stream synchronisation; the device time between the
// first and the last attempt goes to g.last_total_ms.
static int run_attempts(gf_pool* p, int substeps, int use_graph)
{
    if (int r = sync_batch_args(p)) return r;
    if (use_graph) {
        if (p->graph_dirty) {
            if (p->graph_exec) HIPCHK(hipGraphExecDestroy(p->graph_exec));
            if (p->graph) HIPCHK(hipGraphDestroy(p->graph));
            p->graph_exec = nullptr;
            p->graph = nullptr;
            HIPCHK(hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal));
            int r = enqueue_attempt(p, -1);
            hipGraph_t graph = nullptr;
            hipError_t e = hipStreamEndCapture(g.stream, &graph);
            if (r != GF_OK) return r;
            HIPCHK(e);
            p->graph = graph;
            HIPCHK(hipGraphInstantiate(&p->graph_exec, p->graph, nullptr, nullptr, 0));
            p->graph_dirty = false;
        }
        HIPCHK(hipEventRecord(g.ev_begin, g.stream));
        for (int k = 0; k < substeps; k++) HIPCHK(hipGraphLaunch(p->graph_exec, g.stream));
        HIPCHK(hipEventRecord(g.ev_end, g.stream));
        HIPCHK(hipStreamSynchronize(g.stream));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, g.ev_begin, g.ev_end));
        g.last_total_ms = ms;
        g.last_step_ms = 0.0;
        g.last_launches = 0;
    } else {
        if (int r = ensure_step_events(std::max(1, substeps))) return r;
        HIPCHK(hipEventRecord(g.ev_begin, g.stream));
        for (int k = 0; k < substeps; k++)
            if (int r = enqueue_attempt(p, k)) return r;
        HIPCHK(hipEventRecord(g.ev_end, g.stream));
        HIPCHK(hipStreamSynchronize(g.stream));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, g.ev_begin, g.ev_end));
        g.last_total_ms = ms;
        g.last_step_ms = 0.0;
        for (int k = 0; k < substeps; k++) {
            HIPCHK(hipEventElapsedTime(&ms, g.ev_step[2 * k], g.ev_step[2 * k + 1]));
            g.last_step_ms += ms;
        }
        g.last_launches = substeps;
    }
    return GF_OK;
}

int gf_pool_run(gf_pool* p, int substeps, int use_graph, int* any_active)
try {
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    if (substeps < 0) return fail(GF_EINVAL, "substeps must be >= 0");
    if (p->nscen == 0 || p->n == 0) return fail(GF_ESTATE, "no scenario registered");
    if (!p->inflow_error.empty()) return fail(GF_EINVAL, "%s", p->inflow_error.c_str());
    if (int r = run_attempts(p, substeps, use_graph)) return r;
    return any_active_flag(p, any_active);
} GF_CATCH_STATUS

int gf_pool_advance(gf_pool* p, int s, int target_steps, int use_graph, int max_attempts, gf_scenario_info* info)
try {
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    if (s < 0 || s >= p->capacity || p->h_ctrl[s].npatch == 0) return fail(GF_EINVAL, "scenario %d unknown", s);
    if (target_steps <= 0) return fail(GF_EINVAL, "target_steps must be > 0");
    if (!p->inflow_error.empty()) return fail(GF_EINVAL, "%s", p->inflow_error.c_str());
    // the stop condition, as gf_pool_set_scenario_stop sets it (t_end off)
    ScenCtrl& c = p->h_ctrl[s];
    HIPCHK(hipMemcpyAsync(&c, p->ctrl + s, sizeof(ScenCtrl), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    if (max_attempts > 0 && target_steps - c.nsteps > max_attempts)
        return fail(GF_ESTATE, "scenario %d: %d steps short of %d exceed max_attempts %d", s, target_steps - c.nsteps,
                    target_steps, max_attempts);
    c.target_steps = target_steps;
    c.t_end = 0.0;
    c.active = (c.status == gf::SCEN_OK && c.nsteps < target_steps) ? 1 : 0;
    HIPCHK(hipMemcpyAsync(p->ctrl + s, &c, sizeof(ScenCtrl), hipMemcpyHostToDevice, g.stream));
    // one attempt per missing step, then single attempts for the retakes
    double total_ms = 0.0;
    int attempts = 0;
    ScenCtrl now = c;
    while (now.active) {
        const int k = attempts == 0 ? target_steps - now.nsteps : 1;
        if (max_attempts > 0 && attempts + k > max_attempts)
            return fail(GF_ESTATE, "scenario %d: %d steps short of %d after %d attempts", s, target_steps - now.nsteps,
                        target_steps, attempts);
        if (int r = run_attempts(p, k, use_graph)) return r;
        total_ms += g.last_total_ms;
        attempts += k;
        HIPCHK(hipMemcpyAsync(&now, p->ctrl + s, sizeof(ScenCtrl), hipMemcpyDeviceToHost, g.stream));
        HIPCHK(hipStreamSynchronize(g.stream));
    }
    g.last_total_ms = total_ms;
    p->h_ctrl[s] = now;
    if (info) {
        info->t = now.t;
        info->dt = now.dt;
        info->maxcfl = now.maxcfl;
        info->last_dt = now.last_dt;
        info->nsteps = now.nsteps;
        info->retakes = now.retakes;
        info->attempts = now.attempts;
        info->active = now.active;
        info->cur = now.cur;
        info->status = now.status;
    }
    return GF_OK;
} GF_CATCH_STATUS

int gf_pool_run_plain(gf_pool* p, int substeps, int* any_active)
try {
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    if (substeps < 0) return fail(GF_EINVAL, "substeps must be >= 0");
    if (p->nscen == 0 || p->n == 0) return fail(GF_ESTATE, "no scenario registered");
    if (!p->inflow_error.empty()) return fail(GF_EINVAL, "%s", p->inflow_error.c_str());
    std::vector<ScenCtrl>& hc = p->h_ctrl;
    HIPCHK(hipMemcpyAsync(hc.data(), p->ctrl, sizeof(ScenCtrl) * p->nscen, hipMemcpyDeviceToHost, g.stream));
    if (int r = reset_dry_age(p, 0, p->capacity)) return r;
    HIPCHK(hipStreamSynchronize(g.stream));
    double total_ms = 0.0, step_ms = 0.0;
    int launches = 0;
    for (int it = 0; it < substeps; it++) {
        int nact = 0;
        for (int s = 0; s < p->nscen; s++) nact += hc[s].active;
        if (nact == 0) break;
        // patches of active scenarios, one launch each
        int cnt = 0;
        for (int s = 0; s < p->nscen; s++)
            if (hc[s].active) cnt += hc[s].npatch;
        if (int r = ensure_step_events(cnt)) return r;
        HIPCHK(hipEventRecord(g.ev_begin, g.stream));
        if (p->amr) {
            // the multi-level ghost update reads each scenario's buffer and
            // time from the device records: bring them up to date first
            HIPCHK(hipMemcpyAsync(p->ctrl, hc.data(), sizeof(ScenCtrl) * p->nscen, hipMemcpyHostToDevice, g.stream));
            if (int r = enqueue_ghost_update(p, -1, 0.0)) return r;
        }
        int e = 0;
        for (int s = 0; s < p->nscen; s++) {
            if (!hc[s].active) continue;
            gf::GhostArgs ga{p->desc, hc[s].first_patch, hc[s].npatch, p->q[0], p->q[1], nullptr, hc[s].cur};
            HIPCHK(gf::launch_ghost(p->mx, p->my, ga, g.stream));
            if (!p->amr)
                HIPCHK(gf::launch_inflow(p->mx, p->my,
                                         inflow_args(p, hc[s].first_patch, hc[s].first_patch + hc[s].npatch, hc[s].cur,
                                                     hc[s].t),
                                         inflow_grid(), g.stream));
            for (int k = hc[s].first_patch; k < hc[s].first_patch + hc[s].npatch; k++, e++) {
                HIPCHK(hipEventRecord(g.ev_step[2 * e], g.stream));
                HIPCHK(gf::launch_step(p->mx, p->my, p->variant,
                                       step_args(p, k, 1, p->q[0], p->q[1], nullptr, hc[s].dt, hc[s].cur), g.stream));
                HIPCHK(hipEventRecord(g.ev_step[2 * e + 1], g.stream));
            }
        }
        HIPCHK(hipMemcpyAsync(p->h_cfl, p->cfl, sizeof(double) * p->n, hipMemcpyDeviceToHost, g.stream));
        HIPCHK(hipEventRecord(g.ev_end, g.stream));
        HIPCHK(hipStreamSynchronize(g.stream));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, g.ev_begin, g.ev_end));
        total_ms += ms;
        for (int k = 0; k < e; k++) {
            HIPCHK(hipEventElapsedTime(&ms, g.ev_step[2 * k], g.ev_step[2 * k + 1]));
            step_ms += ms;
        }
        launches += e;
        // dt rule on the host (fc2d_geoclaw_run.c:529-565)
        for (int s = 0; s < p->nscen; s++) {
            ScenCtrl& c = hc[s];
            if (!c.active) continue;
            double m = 0.0;
            for (int k = c.first_patch; k < c.first_patch + c.npatch; k++) m = max_nan(m, p->h_cfl[k]);
            c.maxcfl = m;
            c.attempts += 1;
            if (!gf::cfl_finite(m)) {
                c.status = gf::SCEN_NONFINITE;
                c.active = 0;
            } else if (m > c.max_cfl) {
                c.retakes += 1;
                c.dt = c.dt * c.desired_cfl / m;
            } else {
                c.t = c.t + c.dt;
                c.last_dt = c.dt;
                c.cur ^= 1;
                c.nsteps += 1;
                c.dt = gf::next_dt(c.dt, c.desired_cfl, m, c.dt_max);
                if ((c.target_steps > 0 && c.nsteps >= c.target_steps) || (c.t_end > 0.0 && c.t >= c.t_end))
                    c.active = 0;
            }
        }
    }
    HIPCHK(hipMemcpyAsync(p->ctrl, hc.data(), sizeof(ScenCtrl) * p->nscen, hipMemcpyHostToDevice, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    g.last_total_ms = total_ms;
    g.last_step_ms = step_ms;
    g.last_launches = launches;
    if (any_active) {
        int a = 0;
        for (int s = 0; s < p->nscen; s++) a |= hc[s].active;
        *any_active = a;
    }
    return GF_OK;
} GF_CATCH_STATUS

int gf_step2_batch(int mx, int my, int mbc, int maux, int n, double* const* q, double* const* aux,
                   const double* dx, const double* dy, double t, double dt, double* maxcfl)
try {
    (void)t;
    if (int r = need_init()) return r;
    if (n <= 0) return fail(GF_EINVAL, "batch size must be > 0");
    if (!q || !aux || !dx || !dy) return fail(GF_EINVAL, "NULL argument");
    if (maux < 1) return fail(GF_EINVAL, "maux must be >= 1");
    if (!(dt > 0.0)) return fail(GF_EINVAL, "dt must be > 0");
    if (mbc != gf::MBC) return fail(GF_EUNSUPPORTED, "mbc = %d; only mbc = 2 is supported", mbc);
    if (!gf::supported_shape(mx, my)) return fail(GF_EUNSUPPORTED, "patch shape %dx%d not supported", mx, my);
    gf_pool* p = g.batch_pool;
    if (!p || p->mx != mx || p->my != my || p->capacity < n) {
        if (p) gf_pool_destroy(p);
        g.batch_pool = nullptr;
        p = gf_pool_create(mx, my, mbc, n);
        if (!p) return GF_ENOMEM;
        g.batch_pool = p;
    }
    const size_t np2 = p->np2;
    // patch descriptors: only dx, dy matter (ghost cells come from the host)
    for (int k = 0; k < n; k++) {
        PatchDesc& d = p->h_desc[k];
        if (!(dx[k] > 0.0) || !(dy[k] > 0.0)) return fail(GF_EINVAL, "patch %d: dx, dy must be > 0", k);
        d.dx = dx[k];
        d.dy = dy[k];
        d.scen = 0;
        for (int s = 0; s < 8; s++) d.nbr[s] = GF_NB_STORED;
    }
    p->n = n;
    HIPCHK(hipMemcpyAsync(p->desc, p->h_desc.data(), sizeof(PatchDesc) * n, hipMemcpyHostToDevice, g.stream));
    // pack: host patches -> pinned staging -> device SoA in both buffers (the
    // step reads buffer 0 and writes the interior of buffer 1, whose ghost
    // cells keep the host's values), then buffer 1 back the same way.  Never
    // in place: the tiles of a tiled variant read each other's cells.
    if (int r = staged_upload(p, 0, n, 0, maux, true, [&](int k, double* sq, double* sa) {
            std::memcpy(sq, q[k], 3 * np2 * sizeof(double));
            std::memcpy(sa, aux[k], maux * np2 * sizeof(double));
        }, 1))
        return r;
    HIPCHK(gf::launch_step(mx, my, p->variant, step_args(p, 0, n, p->q[0], p->q[1], nullptr, dt, 0), g.stream));
    HIPCHK(gf::launch_max(p->cfl, n, g.d_max, g.stream));
    HIPCHK(hipMemcpyAsync(g.h_max, g.d_max, sizeof(double), hipMemcpyDeviceToHost, g.stream));
    if (int r = staged_download(p, 0, n, 1, [&](int k, const double* sq) {
            std::memcpy(q[k], sq, 3 * np2 * sizeof(double));
        }))
        return r;
    if (maxcfl) *maxcfl = *g.h_max;
    return GF_OK;
} GF_CATCH_STATUS

// Layout of the no-allocation batch in the device scratch array, in doubles,
// each region 256-byte aligned.
struct NoAllocLayout {
    size_t q0, q1, aux, stage, dxy, desc, cfl, tiles, dmax, total;
};
static NoAllocLayout noalloc_layout(int n, int maux, size_t np2)
{
    auto al = [](size_t d) { return (d + 31) / 32 * 32; };
    NoAllocLayout L{};
    size_t off = 0;
    L.q0 = off;    off += al((size_t)n * 3 * np2);
    L.q1 = off;    off += al((size_t)n * 3 * np2);
    L.aux = off;   off += al((size_t)n * np2);
    L.stage = off; off += al((size_t)n * (3 + (size_t)maux) * np2);
    L.dxy = off;   off += al(2 * (size_t)n);
    L.desc = off;  off += al((size_t)n * sizeof(PatchDesc) / sizeof(double));
    L.cfl = off;   off += al((size_t)n);
    L.tiles = off; off += al(2 * (size_t)n + ((size_t)n + 1) / 2);  // tile CFL, depth (8 B), count (4 B)
    L.dmax = off;  off += 32;
    L.total = off;
    return L;
}

int gf_step2_batch_noalloc(int mx, int my, int mbc, int maux, int n, double* const* q, double* const* aux,
                           const double* dx, const double* dy, double t, double dt, double* maxcfl)
try {
    (void)t;
    if (int r = need_init()) return r;
    if (n <= 0) return fail(GF_EINVAL, "batch size must be > 0");
    if (!q || !aux || !dx || !dy) return fail(GF_EINVAL, "NULL argument");
    if (maux < 1) return fail(GF_EINVAL, "maux must be >= 1");
    if (!(dt > 0.0)) return fail(GF_EINVAL, "dt must be > 0");
    if (mbc != gf::MBC) return fail(GF_EUNSUPPORTED, "mbc = %d; only mbc = 2 is supported", mbc);
    if (!gf::supported_shape(mx, my)) return fail(GF_EUNSUPPORTED, "patch shape %dx%d not supported", mx, my);
    for (int k = 0; k < n; k++)
        if (!(dx[k] > 0.0) || !(dy[k] > 0.0)) return fail(GF_EINVAL, "patch %d: dx, dy must be > 0", k);
    double* base = g.device < Lib::kMaxDevices ? g.scratch_dev[g.device] : nullptr;
    if (!base) return fail(GF_ESTATE, "no scratch address on device %d", g.device);
    const size_t np2 = (size_t)(mx + 2 * mbc) * (my + 2 * mbc);
    const NoAllocLayout L = noalloc_layout(n, maux, np2);
    if (L.total * sizeof(double) > kNoAllocBytes)
        return fail(GF_EINVAL, "a batch of %d patches of %dx%d with maux = %d needs %zu bytes of device scratch; "
                    "gf_step2_batch_noalloc holds %zu (use gf_step2_batch, or smaller batches)",
                    n, mx, my, maux, L.total * sizeof(double), kNoAllocBytes);
    double* q0 = base + L.q0;
    double* q1 = base + L.q1;
    double* st = base + L.stage;
    double* dxy = base + L.dxy;
    PatchDesc* desc = reinterpret_cast<PatchDesc*>(base + L.desc);
    unsigned long long* tcfl = reinterpret_cast<unsigned long long*>(base + L.tiles);
    // host patches straight into the staging area (no host staging either)
    const size_t qd = (size_t)n * 3 * np2;
    for (int k = 0; k < n; k++) {
        HIPCHK(hipMemcpyAsync(st + (size_t)k * 3 * np2, q[k], sizeof(double) * 3 * np2, hipMemcpyHostToDevice, g.stream));
        HIPCHK(hipMemcpyAsync(st + qd + (size_t)k * maux * np2, aux[k], sizeof(double) * maux * np2,
                              hipMemcpyHostToDevice, g.stream));
    }
    HIPCHK(hipMemcpyAsync(dxy, dx, sizeof(double) * n, hipMemcpyHostToDevice, g.stream));
    HIPCHK(hipMemcpyAsync(dxy + n, dy, sizeof(double) * n, hipMemcpyHostToDevice, g.stream));
    HIPCHK(gf::launch_desc_fill(desc, dxy, dxy + n, n, g.stream));
    // the tile fold slots start at zero (an earlier call may have stopped half way)
    HIPCHK(gf::launch_zero_words(reinterpret_cast<int*>(tcfl), 2 * (2 * (size_t)n) + (size_t)n, g.stream));
    HIPCHK(gf::launch_pack(st, st + qd, q0, q1, base + L.aux, 0, n, (int)np2, maux, g.stream));
    gf::StepArgs a{};
    a.desc = desc;
    a.first = 0;
    a.count = n;
    a.q0 = q0;
    a.q1 = q1;
    a.aux = base + L.aux;
    a.cfl_out = base + L.cfl;
    a.dt = dt;
    a.cur = 0;
    a.P = g.P;
    a.tile_cfl = tcfl;
    a.tile_h = tcfl + n;
    a.tile_cnt = reinterpret_cast<int*>(tcfl + 2 * (size_t)n);
    HIPCHK(gf::launch_step(mx, my, 0, a, g.stream));
    HIPCHK(gf::launch_max(base + L.cfl, n, base + L.dmax, g.stream));
    HIPCHK(gf::launch_unpack(st, q1, 0, n, (int)np2, g.stream));
    for (int k = 0; k < n; k++)
        HIPCHK(hipMemcpyAsync(q[k], st + (size_t)k * 3 * np2, sizeof(double) * 3 * np2, hipMemcpyDeviceToHost, g.stream));
    double m = 0.0;
    HIPCHK(hipMemcpyAsync(&m, base + L.dmax, sizeof(double), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    if (maxcfl) *maxcfl = m;
    return GF_OK;
} GF_CATCH_STATUS

int gf_eval_powers(int which, int n, const double* x, double* out)
try {
    if (int r = need_init()) return r;
    if (which != 0 && which != 1) return fail(GF_EINVAL, "which must be 0 (7/3) or 1 (2/3)");
    if (n < 0 || (n > 0 && (!x || !out))) return fail(GF_EINVAL, "invalid arrays");
    if (n == 0) return GF_OK;
    double* d = nullptr;
    HIPCHK(hipMalloc(&d, sizeof(double) * 2 * (size_t)n));
    hipError_t e = hipMemcpyAsync(d, x, sizeof(double) * n, hipMemcpyHostToDevice, g.stream);
    if (e == hipSuccess) e = gf::launch_eval_powers(which, d, d + n, n, g.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d + n, sizeof(double) * n, hipMemcpyDeviceToHost, g.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g.stream);
    hipFree(d);
    if (e != hipSuccess) return fail(GF_EHIP, "gf_eval_powers: %s", hipGetErrorString(e));
    return GF_OK;
} GF_CATCH_STATUS

int gf_synchronize(void)
try {
    if (int r = need_init()) return r;
    HIPCHK(hipStreamSynchronize(g.stream));
    return GF_OK;
} GF_CATCH_STATUS

int gf_pool_stats(gf_pool* p, long long* active_patches, long long* dry_patches, long long* attempts,
                  long long* steady_patches)
try {
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    long long s[4];
    HIPCHK(hipMemcpyAsync(s, p->stats, sizeof(s), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    if (active_patches) *active_patches = s[0];
    if (dry_patches) *dry_patches = s[1];
    if (attempts) *attempts = s[2];
    if (steady_patches) *steady_patches = s[3];
    return GF_OK;
} GF_CATCH_STATUS

static int tag_impl(gf_pool* p, bool coarsen, int buf, double t, const gf_refine* c, int level, int maxlevel,
                    int init_flag, const int* patches, int n, int* tags)
{
    if (!p || !c) return fail(GF_EINVAL, "pool or criteria is NULL");
    if (buf != 0 && buf != 1) return fail(GF_EINVAL, "buffer must be 0 or 1");
    if (n < 0 || (n > 0 && (!patches || !tags))) return fail(GF_EINVAL, "bad patch list");
    if (c->num_speeds < 1 || !c->speed_tolerance) return fail(GF_EINVAL, "speed_tolerance needs >= 1 entry");
    if (c->num_flowgrades < 0 || (c->num_flowgrades > 0 && (!c->flowgradevalue || !c->iflowgradevariable ||
                                                            !c->iflowgrademinlevel)))
        return fail(GF_EINVAL, "flowgrade arrays missing");
    if (c->mregions < 0 || (c->mregions > 0 && !c->region)) return fail(GF_EINVAL, "region array missing");
    if (n == 0) return GF_OK;
    const int np = coarsen ? 4 * n : n;
    for (int k = 0; k < np; k++)
        if (patches[k] < 0 || patches[k] >= p->capacity) return fail(GF_EINVAL, "patch %d outside the pool", patches[k]);
    // one allocation: criteria struct, its arrays, the patch list, the tags
    const int nfg = c->num_flowgrades, nreg = c->mregions;
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t o_sp = al(sizeof(gf::TagCriteria));
    const size_t o_fv = al(o_sp + sizeof(double) * c->num_speeds);
    const size_t o_rg = al(o_fv + sizeof(double) * nfg);
    const size_t o_var = al(o_rg + sizeof(double) * 8 * nreg);
    const size_t o_min = al(o_var + sizeof(int) * nfg);
    const size_t o_pat = al(o_min + sizeof(int) * nfg);
    const size_t o_out = al(o_pat + sizeof(int) * np);
    const size_t bytes = al(o_out + sizeof(int) * n);
    if (bytes > p->tag_bytes) {
        if (p->tag_buf) HIPCHK(hipFree(p->tag_buf));
        p->tag_buf = nullptr;
        p->tag_bytes = 0;
        HIPCHK(hipMalloc(&p->tag_buf, bytes));
        p->tag_bytes = bytes;
    }
    char* d = (char*)p->tag_buf;
    std::vector<char> h(bytes, 0);
    gf::TagCriteria tc{};
    tc.wave_tolerance = c->wave_tolerance;
    tc.deep_depth = c->deep_depth;
    tc.max_velocity_depth_product = c->max_velocity_depth_product;
    tc.sea_level = c->sea_level;
    tc.max_level_deep = c->max_level_deep;
    tc.num_speeds = c->num_speeds;
    tc.num_flowgrades = nfg;
    tc.mregions = nreg;
    tc.speed_tolerance = (const double*)(d + o_sp);
    tc.fg_value = (const double*)(d + o_fv);
    tc.region = (const double*)(d + o_rg);
    tc.fg_var = (const int*)(d + o_var);
    tc.fg_minlevel = (const int*)(d + o_min);
    std::memcpy(h.data(), &tc, sizeof(tc));
    std::memcpy(h.data() + o_sp, c->speed_tolerance, sizeof(double) * c->num_speeds);
    if (nfg) {
        std::memcpy(h.data() + o_fv, c->flowgradevalue, sizeof(double) * nfg);
        std::memcpy(h.data() + o_var, c->iflowgradevariable, sizeof(int) * nfg);
        std::memcpy(h.data() + o_min, c->iflowgrademinlevel, sizeof(int) * nfg);
    }
    if (nreg) std::memcpy(h.data() + o_rg, c->region, sizeof(double) * 8 * nreg);
    std::memcpy(h.data() + o_pat, patches, sizeof(int) * np);
    HIPCHK(hipMemcpyAsync(d, h.data(), o_out, hipMemcpyHostToDevice, g.stream));
    gf::TagArgs a{};
    a.desc = p->desc;
    a.q = p->q[buf];
    a.aux = p->aux;
    a.crit = (const gf::TagCriteria*)d;
    a.patches = (const int*)(d + o_pat);
    a.n = n;
    a.out = (int*)(d + o_out);
    a.t = t;
    a.drytol = g.P.dry_tolerance;
    a.level = level;
    a.maxlevel = maxlevel;
    a.init_flag = init_flag;
    HIPCHK(gf::launch_tag(p->mx, p->my, coarsen, a, g.stream));
    HIPCHK(hipMemcpyAsync(tags, d + o_out, sizeof(int) * n, hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    return GF_OK;
}

int gf_pool_tag_refine(gf_pool* p, int buf, double t, const gf_refine* crit, int level, int maxlevel,
                       int init_flag, const int* patches, int n, int* tags)
try {
    return tag_impl(p, false, buf, t, crit, level, maxlevel, init_flag, patches, n, tags);
} GF_CATCH_STATUS

int gf_pool_tag_coarsen(gf_pool* p, int buf, double t, const gf_refine* crit, int level, int maxlevel,
                        int init_flag, const int* families, int n, int* tags)
try {
    return tag_impl(p, true, buf, t, crit, level, maxlevel, init_flag, families, n, tags);
} GF_CATCH_STATUS

int gf_pool_variants(gf_pool* p, int* current, int* count, int* desc)
try {
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    const int n = gf::num_variants(p->mx, p->my);
    if (current) *current = p->variant;
    if (count) *count = n;
    if (desc)
        for (int k = 0; k < n; k++) {
            gf::StepVariant v;
            gf::get_variant(p->mx, p->my, k, &v);
            desc[5 * k] = v.tx;
            desc[5 * k + 1] = v.ty;
            desc[5 * k + 2] = v.nt;
            desc[5 * k + 3] = v.npp;
            desc[5 * k + 4] = v.freg;
        }
    return GF_OK;
} GF_CATCH_STATUS

int gf_pool_set_variant(gf_pool* p, int variant)
try {
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    if (variant < 0 || variant >= gf::num_variants(p->mx, p->my))
        return fail(GF_EINVAL, "variant %d outside [0,%d)", variant, gf::num_variants(p->mx, p->my));
    if (variant != p->variant) {
        p->variant = variant;
        p->wet_grid = 0;
        p->graph_dirty = true;
    }
    return GF_OK;
} GF_CATCH_STATUS

int gf_pool_autotune(gf_pool* p, int reps, int* best, double* ms)
try {
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    if (reps < 1) return fail(GF_EINVAL, "reps must be >= 1");
    if (p->nscen == 0 || p->n == 0) return fail(GF_ESTATE, "no scenario registered");
    if (!p->inflow_error.empty()) return fail(GF_EINVAL, "%s", p->inflow_error.c_str());
    if (int r = sync_batch_args(p)) return r;
    const int nv = gf::num_variants(p->mx, p->my);
    // The active list of the current states (what the next attempt steps).
    // Each candidate steps it from the accepted buffer into the other one,
    // which the next attempt overwrites: no scenario advances.
    if (p->amr) {
        if (int r = enqueue_ghost_update(p, -1, 0.0)) return r;
    } else {
        HIPCHK(gf::launch_inflow(p->mx, p->my, inflow_args(p, 0, p->capacity, -1, 0.0), inflow_grid(), g.stream));
    }
    HIPCHK(gf::launch_classify(p->d_args, p->capacity, g.stream));
    // an inactive scenario (stopped at its target, or between regrids) lists
    // nothing: with every list empty there is no work to time
    int cnt[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(cnt, p->counters + gf::CNT_WET, sizeof(cnt), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    if (cnt[0] == 0 && cnt[1] == 0) {
        HIPCHK(hipMemsetAsync(p->counters, 0, gf::NCOUNTERS * sizeof(int), g.stream));
        HIPCHK(hipStreamSynchronize(g.stream));
        return fail(GF_ESTATE, "no active or dry patch to time: every scenario is inactive or steady");
    }
    if (int r = ensure_pair_events(g.ev_tune, 1)) return r;
    const int saved = p->variant;
    double best_ms = 0.0;
    int best_v = saved;
    for (int v = 0; v < nv; v++) {
        p->variant = v;
        p->wet_grid = 0;
        double tot = 0.0;
        for (int r = 0; r <= reps; r++) {  // the first run warms up
            HIPCHK(hipMemsetAsync(p->counters + gf::NEXT_WET, 0, sizeof(int), g.stream));
            HIPCHK(hipEventRecord(g.ev_tune[0], g.stream));
            HIPCHK(gf::launch_all(p->mx, p->my, v, gf::default_method(g.P), p->d_args, wet_grid(p), g.stream));
            HIPCHK(hipEventRecord(g.ev_tune[1], g.stream));
            HIPCHK(hipEventSynchronize(g.ev_tune[1]));
            float t = 0.f;
            HIPCHK(hipEventElapsedTime(&t, g.ev_tune[0], g.ev_tune[1]));
            if (r > 0) tot += t;
        }
        const double avg = tot / reps;
        if (ms) ms[v] = avg;
        if (v == 0 || avg < best_ms) {
            best_ms = avg;
            best_v = v;
        }
    }
    // undo the candidates' side effects on the attempt's counters and maxima
    HIPCHK(hipMemsetAsync(p->counters, 0, gf::NCOUNTERS * sizeof(int), g.stream));
    HIPCHK(hipMemsetAsync(p->cflbits, 0, (size_t)p->capacity * sizeof(unsigned long long), g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    p->variant = saved;
    p->wet_grid = 0;
    if (best) *best = best_v;
    return gf_pool_set_variant(p, best_v);
} GF_CATCH_STATUS

int gf_last_timing(double* total_ms, double* step_ms, int* step_launches)
try {
    if (total_ms) *total_ms = g.last_total_ms;
    if (step_ms) *step_ms = g.last_step_ms;
    if (step_launches) *step_launches = g.last_launches;
    return GF_OK;
} GF_CATCH_STATUS

// ---- output: ASCII frames and gauges ----------------------------------------

// Per-pool device scratch of the tagging and gauge calls, grown on demand.
static int ensure_scratch(gf_pool* p, size_t bytes)
{
    if (bytes <= p->tag_bytes) return GF_OK;
    if (p->tag_buf) HIPCHK(hipFree(p->tag_buf));
    p->tag_buf = nullptr;
    p->tag_bytes = 0;
    HIPCHK(hipMalloc(&p->tag_buf, bytes));
    p->tag_bytes = bytes;
    return GF_OK;
}

static std::string join_path(const char* dir, const char* name)
{
    std::string s(dir);
    if (!s.empty() && s.back() != '/') s.push_back('/');
    return s + name;
}

int gf_pool_write_frame(gf_pool* p, int buf, int first, int n, const char* dir, int frame, double time,
                        int ngrids, int write_header, const int* patch_num, const int* level, const int* blockno,
                        int mpirank)
try {
    if (!p || !dir) return fail(GF_EINVAL, "pool or dir is NULL");
    if (buf != 0 && buf != 1) return fail(GF_EINVAL, "buffer must be 0 or 1");
    if (first < 0 || n < 0 || first + n > p->capacity) return fail(GF_EINVAL, "slots [%d,%d) outside the pool", first, first + n);
    if (frame < 0) return fail(GF_EINVAL, "frame must be >= 0");
    for (int k = 0; k < n; k++)
        if (!(p->h_desc[first + k].dx > 0.0)) return fail(GF_EINVAL, "slot %d is not described", first + k);
    char qname[32], tname[32];
    std::snprintf(qname, sizeof qname, "fort.q%04d", frame % 10000);
    std::snprintf(tname, sizeof tname, "fort.t%04d", frame % 10000);
    const std::string qpath = join_path(dir, qname);
    if (write_header) {
        std::string h;
        gf::ascii::frame_header(h, time, 3, 1, ngrids, p->mbc);
        FILE* f = std::fopen(join_path(dir, tname).c_str(), "wb");
        if (!f) return fail(GF_EIO, "cannot write %s", join_path(dir, tname).c_str());
        const bool ok = std::fwrite(h.data(), 1, h.size(), f) == h.size();
        if (std::fclose(f) != 0 || !ok) return fail(GF_EIO, "cannot write %s", tname);
        f = std::fopen(qpath.c_str(), "wb");
        if (!f || std::fclose(f) != 0) return fail(GF_EIO, "cannot write %s", qpath.c_str());
    }
    FILE* fq = std::fopen(qpath.c_str(), "ab");
    if (!fq) return fail(GF_EIO, "cannot write %s", qpath.c_str());
    const int cells = p->mx * p->my;
    const size_t chunk = std::min<size_t>(kChunk, (size_t)p->capacity);  // d_stage holds >= 4 * cells per patch
    std::vector<std::string> text(chunk);
    int rc = GF_OK;
    for (int off = 0; off < n && rc == GF_OK; off += (int)chunk) {
        const int m = std::min<int>((int)chunk, n - off);
        hipError_t e = gf::launch_frame_pack(p->mx, p->my, p->q[buf], p->aux, first + off, m, p->d_stage, g.stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(p->h_stage, p->d_stage, sizeof(double) * (size_t)m * cells * 4, hipMemcpyDeviceToHost,
                               g.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(g.stream);
        if (e != hipSuccess) {
            rc = fail(GF_EHIP, "frame pack: %s", hipGetErrorString(e));
            break;
        }
        // patches formatted on the library's host threads, written in order
        host_parallel(m, [&](int k) {
            const int s = first + off + k;
            const PatchDesc& d = p->h_desc[s];
            std::string& out = text[k];
            out.clear();
            out.reserve(400 + (size_t)cells * 105 + (size_t)p->my * 3);
            gf::ascii::frame_patch(out, p->h_stage + (size_t)k * cells * 4, p->mx, p->my, d.xlo, d.ylo, d.dx, d.dy,
                                   patch_num ? patch_num[off + k] : s, level ? level[off + k] : 0,
                                   blockno ? blockno[off + k] : 0, mpirank);
        });
        for (int k = 0; k < m; k++)
            if (std::fwrite(text[k].data(), 1, text[k].size(), fq) != text[k].size()) {
                rc = fail(GF_EIO, "cannot write %s", qpath.c_str());
                break;
            }
    }
    if (std::fclose(fq) != 0 && rc == GF_OK) rc = fail(GF_EIO, "cannot write %s", qpath.c_str());
    // the flush below 1e-99 may have changed the state
    if (int r = reset_dry_age(p, first, n)) return r;
    HIPCHK(hipStreamSynchronize(g.stream));
    return rc;
} GF_CATCH_STATUS

int gf_pool_sample_gauges(gf_pool* p, int buf, int n, const int* slot, const double* xc, const double* yc,
                          double* out)
try {
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    if (buf != 0 && buf != 1) return fail(GF_EINVAL, "buffer must be 0 or 1");
    if (n < 0) return fail(GF_EINVAL, "n must be >= 0");
    if (n == 0) return GF_OK;
    if (!slot || !xc || !yc || !out) return fail(GF_EINVAL, "NULL argument");
    for (int k = 0; k < n; k++) {
        if (slot[k] < 0 || slot[k] >= p->capacity) return fail(GF_EINVAL, "gauge %d: slot %d outside the pool", k, slot[k]);
        const PatchDesc& d = p->h_desc[slot[k]];
        if (!(d.dx > 0.0)) return fail(GF_EINVAL, "gauge %d: slot %d is not described", k, slot[k]);
        if (!(xc[k] >= d.xlo && xc[k] <= d.xlo + p->mx * d.dx && yc[k] >= d.ylo && yc[k] <= d.ylo + p->my * d.dy))
            return fail(GF_EINVAL, "gauge %d at (%g, %g) is outside patch %d", k, xc[k], yc[k], slot[k]);
    }
    const size_t o_xy = ((sizeof(int) * (size_t)n + 255) / 256) * 256;
    const size_t o_out = o_xy + sizeof(double) * 2 * (size_t)n;
    const size_t bytes = o_out + sizeof(double) * 4 * (size_t)n;
    if (int r = ensure_scratch(p, bytes)) return r;
    std::vector<double> xy(2 * (size_t)n);
    for (int k = 0; k < n; k++) {
        xy[2 * k] = xc[k];
        xy[2 * k + 1] = yc[k];
    }
    char* d = (char*)p->tag_buf;
    HIPCHK(hipMemcpyAsync(d, slot, sizeof(int) * n, hipMemcpyHostToDevice, g.stream));
    HIPCHK(hipMemcpyAsync(d + o_xy, xy.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice, g.stream));
    gf::GaugeArgs a{};
    a.desc = p->desc;
    a.q = p->q[buf];
    a.aux = p->aux;
    a.slot = (const int*)d;
    a.xy = (const double*)(d + o_xy);
    a.out = (double*)(d + o_out);
    a.n = n;
    a.dry_tolerance = g.P.dry_tolerance;
    HIPCHK(gf::launch_gauges(p->mx, p->my, a, g.stream));
    HIPCHK(hipMemcpyAsync(out, d + o_out, sizeof(double) * 4 * n, hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    return GF_OK;
} GF_CATCH_STATUS

// ---- regrid and time interpolation between levels ----------------------------

// The family lists in the pool's device scratch: n coarse slots, then 4n
// child slots.
static int regrid_lists(gf_pool* p, int n, const int* cs, const int* fs, int** dcs, int** dfs)
{
    const size_t o_fs = ((sizeof(int) * (size_t)n + 255) / 256) * 256;
    const size_t bytes = o_fs + sizeof(int) * 4 * (size_t)n;
    if (int r = ensure_scratch(p, bytes)) return r;
    char* d = (char*)p->tag_buf;
    HIPCHK(hipMemcpyAsync(d, cs, sizeof(int) * n, hipMemcpyHostToDevice, g.stream));
    HIPCHK(hipMemcpyAsync(d + o_fs, fs, sizeof(int) * 4 * n, hipMemcpyHostToDevice, g.stream));
    *dcs = (int*)d;
    *dfs = (int*)(d + o_fs);
    return GF_OK;
}

// Slots must be described, distinct, and (for interpolation) no child may be
// a coarse patch of the call or a same-level neighbour that a coarse patch's
// halo reads.
static int regrid_check(gf_pool* p, int buf, int n, const int* cs, const int* fs, bool interp)
{
    if (buf != 0 && buf != 1) return fail(GF_EINVAL, "buffer must be 0 or 1");
    if (n < 0 || (n > 0 && (!cs || !fs))) return fail(GF_EINVAL, "invalid family lists");
    if (p->mx % 2 || p->my % 2) return fail(GF_EUNSUPPORTED, "patch size must be even");
    std::vector<char> role(p->capacity, 0);  // 1 coarse, 2 fine, 3 read by a coarse halo
    for (int f = 0; f < n; f++) {
        const int c = cs[f];
        if (c < 0 || c >= p->capacity || !(p->h_desc[c].dx > 0.0)) return fail(GF_EINVAL, "family %d: coarse slot %d invalid", f, c);
        if (role[c]) return fail(GF_EINVAL, "slot %d appears twice", c);
        role[c] = 1;
    }
    for (int f = 0; f < n && interp; f++)
        for (int s = 0; s < 8; s++) {
            const int nb = p->h_desc[cs[f]].nbr[s];
            if (nb >= 0 && !role[nb]) role[nb] = 3;
        }
    for (int k = 0; k < 4 * n; k++) {
        const int c = fs[k];
        if (c < 0 || c >= p->capacity || !(p->h_desc[c].dx > 0.0)) return fail(GF_EINVAL, "child %d: slot %d invalid", k, c);
        if (role[c] == 3) return fail(GF_EINVAL, "child slot %d is read by a coarse patch's halo", c);
        if (role[c]) return fail(GF_EINVAL, "slot %d appears twice", c);
        role[c] = 2;
    }
    return GF_OK;
}

static int regrid_impl(gf_pool* p, int buf, int n, const int* cs, const int* fs, bool interp)
{
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    if (int r = regrid_check(p, buf, n, cs, fs, interp)) return r;
    if (n == 0) return GF_OK;
    int *dcs = nullptr, *dfs = nullptr;
    if (int r = regrid_lists(p, n, cs, fs, &dcs, &dfs)) return r;
    gf::RegridArgs a{};
    a.desc = p->desc;
    a.q = p->q[buf];
    a.aux = p->aux;
    a.cs = dcs;
    a.fs = dfs;
    a.n = n;
    a.dry_tolerance = g.P.dry_tolerance;
    double* hm = p->hmax2 + (size_t)buf * p->capacity;
    if (interp) {
        HIPCHK(gf::launch_interpolate2fine(p->mx, p->my, a, g.stream));
        HIPCHK(gf::launch_hmax_list(p->mx, p->my, p->q[buf], hm, p->dry_age, dfs, 4 * n, g.stream));
    } else {
        HIPCHK(gf::launch_average2coarse(p->mx, p->my, a, g.stream));
        HIPCHK(gf::launch_hmax_list(p->mx, p->my, p->q[buf], hm, p->dry_age, dcs, n, g.stream));
    }
    HIPCHK(hipStreamSynchronize(g.stream));
    return GF_OK;
}

int gf_pool_average2coarse(gf_pool* p, int buf, int n, const int* coarse, const int* fine)
try {
    return regrid_impl(p, buf, n, coarse, fine, false);
} GF_CATCH_STATUS

int gf_pool_interpolate2fine(gf_pool* p, int buf, int n, const int* coarse, const int* fine)
try {
    return regrid_impl(p, buf, n, coarse, fine, true);
} GF_CATCH_STATUS

// Links of a coarse/fine ghost transfer: slots described and distinct per
// link, side 0..7, igrid 0/1 on faces; each written region named once (the
// coarse patch's for averaging, the fine patch's for interpolation); for
// interpolation no slot may be both a coarse and a fine patch of the call,
// since the coarse reads include the coarse patch's stored ghost cells.
static int cf_ghost_check(gf_pool* p, int buf, int n, const int* cs, const int* fs, const int* side,
                          const int* igrid, bool interp)
{
    if (buf != 0 && buf != 1) return fail(GF_EINVAL, "buffer must be 0 or 1");
    if (n < 0 || (n > 0 && (!cs || !fs || !side || !igrid))) return fail(GF_EINVAL, "invalid link lists");
    if (p->mx != p->my || p->mx % 2) return fail(GF_EUNSUPPORTED, "coarse/fine ghost transfers need square patches of even size");
    std::vector<char> role(p->capacity, 0);          // interpolation: 1 coarse, 2 fine
    std::vector<uint16_t> seen(p->capacity, 0);      // regions written: 4 faces x 2 halves, 4 corners
    for (int l = 0; l < n; l++) {
        const int c = cs[l], f = fs[l], sd = side[l];
        if (c < 0 || c >= p->capacity || !(p->h_desc[c].dx > 0.0)) return fail(GF_EINVAL, "link %d: coarse slot %d invalid", l, c);
        if (f < 0 || f >= p->capacity || !(p->h_desc[f].dx > 0.0)) return fail(GF_EINVAL, "link %d: fine slot %d invalid", l, f);
        if (c == f) return fail(GF_EINVAL, "link %d: coarse and fine slot are the same", l);
        if (sd < 0 || sd > 7) return fail(GF_EINVAL, "link %d: side %d not in 0..7", l, sd);
        if (sd < 4 && igrid[l] != 0 && igrid[l] != 1) return fail(GF_EINVAL, "link %d: igrid %d not 0 or 1", l, igrid[l]);
        const uint16_t bit = (uint16_t)(1u << (sd < 4 ? 2 * sd + igrid[l] : 8 + (sd - 4)));
        const int w = interp ? f : c;
        if (seen[w] & bit) return fail(GF_EINVAL, "link %d: slot %d's region is written twice", l, w);
        seen[w] |= bit;
        if (interp) {
            if (role[c] == 2 || role[f] == 1) return fail(GF_EINVAL, "link %d: a slot is both a coarse and a fine patch of the call", l);
            role[c] = 1;
            role[f] = 2;
        }
    }
    return GF_OK;
}

static int cf_ghost_impl(gf_pool* p, int buf, int n, const int* cs, const int* fs, const int* side,
                         const int* igrid, bool interp)
{
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    if (int r = cf_ghost_check(p, buf, n, cs, fs, side, igrid, interp)) return r;
    if (n == 0) return GF_OK;
    // device copies of the four lists, each 256-byte aligned, in the pool's scratch
    const size_t stride = ((sizeof(int) * (size_t)n + 255) / 256) * 256;
    if (int r = ensure_scratch(p, 4 * stride)) return r;
    char* d = (char*)p->tag_buf;
    const int* lists[4] = {cs, fs, side, igrid};
    std::vector<int> ig0;
    for (int k = 0; k < 4; k++) {
        const int* src = lists[k];
        if (k == 3) {  // corners: igrid is not read, pass 0
            ig0.assign(igrid, igrid + n);
            for (int l = 0; l < n; l++)
                if (side[l] >= 4) ig0[l] = 0;
            src = ig0.data();
        }
        HIPCHK(hipMemcpyAsync(d + k * stride, src, sizeof(int) * n, hipMemcpyHostToDevice, g.stream));
    }
    gf::RegridArgs a{};
    a.desc = p->desc;
    a.q = p->q[buf];
    a.aux = p->aux;
    a.cs = (const int*)d;
    a.fs = (const int*)(d + stride);
    a.side = (const int*)(d + 2 * stride);
    a.igrid = (const int*)(d + 3 * stride);
    a.n = n;
    a.dry_tolerance = g.P.dry_tolerance;
    HIPCHK(interp ? gf::launch_interpolate_ghost(p->mx, p->my, a, g.stream)
                  : gf::launch_average_ghost(p->mx, p->my, a, g.stream));
    // the written patches' ghost cells changed: none of them may be skipped
    // as steady-dry on the strength of its old neighbourhood (the interior
    // depth maxima are recomputed unchanged)
    HIPCHK(gf::launch_hmax_list(p->mx, p->my, p->q[buf], p->hmax2 + (size_t)buf * p->capacity, p->dry_age,
                                interp ? a.fs : a.cs, n, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    return GF_OK;
}

int gf_pool_average_ghost(gf_pool* p, int buf, int n, const int* coarse, const int* fine, const int* side,
                          const int* igrid)
try {
    return cf_ghost_impl(p, buf, n, coarse, fine, side, igrid, false);
} GF_CATCH_STATUS

int gf_pool_interpolate_ghost(gf_pool* p, int buf, int n, const int* coarse, const int* fine, const int* side,
                              const int* igrid)
try {
    return cf_ghost_impl(p, buf, n, coarse, fine, side, igrid, true);
} GF_CATCH_STATUS

// ---- resident AMR ---------------------------------------------------------

static int ensure_amr(gf_pool* p)
{
    if (!p->d_ranges) {
        HIPCHK(hipMalloc(&p->d_ranges, sizeof(int) * kRanges));
        HIPCHK(hipMemsetAsync(p->d_ranges, 0, sizeof(int) * kRanges, g.stream));
        HIPCHK(hipStreamSynchronize(g.stream));
    }
    if (!p->amr) {
        p->amr = true;
        p->graph_dirty = true;
    }
    return GF_OK;
}

int gf_pool_set_levels(gf_pool* p, int first, int n, const int* level)
try {
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    if (first < 0 || n < 0 || first + n > p->capacity) return fail(GF_EINVAL, "slots [%d,%d) outside the pool", first, first + n);
    if (n > 0 && !level) return fail(GF_EINVAL, "level is NULL");
    for (int k = 0; k < n; k++)
        if (level[k] < 0 || level[k] >= kMaxLevels) return fail(GF_EINVAL, "slot %d: level %d not in [0, %d)", first + k, level[k], kMaxLevels);
    const bool was_amr = p->amr;
    if (int r = ensure_amr(p)) return r;
    bool in_scenario = false;
    for (int k = 0; k < n; k++) {
        in_scenario |= p->h_desc[first + k].scen >= 0;
        p->h_level[first + k] = level[k];
    }
    const int nlev = 1 + *std::max_element(p->h_level.begin(), p->h_level.end());
    if (nlev != p->nlev) {
        p->nlev = nlev;
        p->graph_dirty = true;
    }
    // the hydrograph list holds the slots of scenarios, by level: levels of
    // slots outside every scenario (a regrid's next half) leave it as it is,
    // once the pool keeps it by level
    return in_scenario || !was_amr ? rebuild_inflow(p) : GF_OK;
} GF_CATCH_STATUS

int gf_pool_set_cf_links(gf_pool* p, int n, const int* coarse, const int* fine, const int* side, const int* igrid)
try {
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    // the same checks as the eager transfers (both directions): slots, sides,
    // each written region once; and a fine patch one level below its coarse one
    if (int r = cf_ghost_check(p, 0, n, coarse, fine, side, igrid, false)) return r;
    {
        std::vector<uint16_t> seen(p->capacity, 0);
        for (int l = 0; l < n; l++) {
            const int sd = side[l];
            const uint16_t bit = (uint16_t)(1u << (sd < 4 ? 2 * sd + igrid[l] : 8 + (sd - 4)));
            if (seen[fine[l]] & bit) return fail(GF_EINVAL, "link %d: slot %d's region is written twice", l, fine[l]);
            seen[fine[l]] |= bit;
            if (p->h_level[fine[l]] != p->h_level[coarse[l]] + 1)
                return fail(GF_EINVAL, "link %d: fine slot %d (level %d) is not one level below coarse slot %d (level %d)", l,
                            fine[l], p->h_level[fine[l]], coarse[l], p->h_level[coarse[l]]);
        }
    }
    if (int r = ensure_amr(p)) return r;
    if (n > p->link_cap) {
        const int cap = std::max(n, 2 * p->link_cap);
        int* d = nullptr;
        HIPCHK(hipMalloc(&d, sizeof(int) * 4 * (size_t)cap));
        HIPCHK(hipStreamSynchronize(g.stream));
        HIPCHK(hipFree(p->d_links));
        p->d_links = d;
        p->graph_dirty = true;  // the graph holds the link arrays' addresses (or had no transfers)
        p->link_cap = cap;
    }
    // links by the coarse patch's level (stable), so each level's
    // interpolation links form one range
    std::vector<int> order(n);
    for (int l = 0; l < n; l++) order[l] = l;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return p->h_level[coarse[a]] < p->h_level[coarse[b]]; });
    std::vector<int> h(4 * (size_t)std::max(p->link_cap, 1), 0);
    for (int k = 0; k < n; k++) {
        const int l = order[k];
        h[k] = coarse[l];
        h[p->link_cap + k] = fine[l];
        h[2 * p->link_cap + k] = side[l];
        h[3 * p->link_cap + k] = side[l] >= 4 ? 0 : igrid[l];
    }
    std::vector<int> rng(2 * (1 + kMaxLevels), 0);
    rng[0] = 0;
    rng[1] = n;
    unsigned levels = 0;
    for (int lv = 0, k = 0; lv < kMaxLevels; lv++) {
        rng[range_interp(lv)] = k;
        while (k < n && p->h_level[h[k]] == lv) k++;
        rng[range_interp(lv) + 1] = k;
        if (k > rng[range_interp(lv)]) levels |= 1u << lv;
    }
    if (levels != p->interp_levels || (n > 0) != (p->nlinks > 0)) {
        p->interp_levels = levels;
        p->graph_dirty = true;
    }
    if (p->link_cap > 0)
        HIPCHK(hipMemcpyAsync(p->d_links, h.data(), sizeof(int) * 4 * (size_t)p->link_cap, hipMemcpyHostToDevice, g.stream));
    HIPCHK(hipMemcpyAsync(p->d_ranges, rng.data(), sizeof(int) * rng.size(), hipMemcpyHostToDevice, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    p->nlinks = n;
    return GF_OK;
} GF_CATCH_STATUS

int gf_pool_ghost_update(gf_pool* p, int buf, double t)
try {
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    if (buf < -1 || buf > 1) return fail(GF_EINVAL, "buffer must be 0, 1 or -1 (each scenario's current one)");
    if (!p->inflow_error.empty()) return fail(GF_EINVAL, "%s", p->inflow_error.c_str());
    if (int r = ensure_amr(p)) return r;
    // stream-ordered: every call that reads the pool afterwards waits for it
    return enqueue_ghost_update(p, buf, t);
} GF_CATCH_STATUS

int gf_pool_move(gf_pool* p, int n, const int* src, const int* dst, int buf)
try {
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    if (buf != 0 && buf != 1) return fail(GF_EINVAL, "buffer must be 0 or 1");
    if (n < 0 || (n > 0 && (!src || !dst))) return fail(GF_EINVAL, "invalid slot lists");
    std::vector<char> role(p->capacity, 0);  // 1 source, 2 destination
    for (int k = 0; k < n; k++) {
        if (src[k] < 0 || src[k] >= p->capacity || dst[k] < 0 || dst[k] >= p->capacity)
            return fail(GF_EINVAL, "entry %d: slot outside the pool", k);
        if (role[dst[k]]) return fail(GF_EINVAL, "entry %d: slot %d is already a source or destination", k, dst[k]);
        role[dst[k]] = 2;
    }
    for (int k = 0; k < n; k++) {
        if (role[src[k]] == 2) return fail(GF_EINVAL, "entry %d: slot %d is both a source and a destination", k, src[k]);
        role[src[k]] = 1;
    }
    if (n == 0) return GF_OK;
    const size_t stride = ((sizeof(int) * (size_t)n + 255) / 256) * 256;
    if (int r = ensure_scratch(p, 2 * stride)) return r;
    char* d = (char*)p->tag_buf;
    HIPCHK(hipMemcpyAsync(d, src, sizeof(int) * n, hipMemcpyHostToDevice, g.stream));
    HIPCHK(hipMemcpyAsync(d + stride, dst, sizeof(int) * n, hipMemcpyHostToDevice, g.stream));
    HIPCHK(gf::launch_move_slots(p->q[buf], p->aux, (const int*)d, (const int*)(d + stride), n, (int)p->np2, g.stream));
    HIPCHK(gf::launch_hmax_list(p->mx, p->my, p->q[buf], p->hmax2 + (size_t)buf * p->capacity, p->dry_age,
                                (const int*)(d + stride), n, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    return GF_OK;
} GF_CATCH_STATUS

int gf_pool_set_scenario_slots(gf_pool* p, int s, int first, int n)
try {
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    if (s < 0 || s >= p->capacity || p->h_ctrl[s].npatch == 0) return fail(GF_EINVAL, "scenario %d unknown", s);
    if (first < 0 || n <= 0 || first + n > p->n)
        return fail(GF_EINVAL, "slots [%d,%d) outside the described slots [0,%d)", first, first + n, p->n);
    ScenCtrl c;
    HIPCHK(hipMemcpyAsync(&c, p->ctrl + s, sizeof(ScenCtrl), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    const int of = c.first_patch, on = c.npatch;
    for (int k = first; k < first + n; k++)
        if (p->h_desc[k].scen >= 0 && p->h_desc[k].scen != s)
            return fail(GF_EINVAL, "slot %d belongs to scenario %d", k, p->h_desc[k].scen);
    for (int k = of; k < of + on; k++) p->h_desc[k].scen = -1;
    for (int k = first; k < first + n; k++) p->h_desc[k].scen = s;
    c.first_patch = first;
    c.npatch = n;
    p->h_ctrl[s] = c;
    // only the scen field changes: set it on the device (the old range
    // first, so slots in both ranges end with s, as on the host)
    HIPCHK(gf::launch_set_scen(p->desc, of, on, -1, g.stream));
    HIPCHK(gf::launch_set_scen(p->desc, first, n, s, g.stream));
    HIPCHK(hipMemcpyAsync(p->ctrl + s, &c, sizeof(ScenCtrl), hipMemcpyHostToDevice, g.stream));
    if (int r = reset_dry_age(p, first, n)) return r;
    HIPCHK(hipStreamSynchronize(g.stream));
    return rebuild_inflow(p);
} GF_CATCH_STATUS

int gf_pool_tag_eta_gradient(gf_pool* p, int buf, double eta_tol, const int* patches, int n, int* tags)
try {
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    if (buf != 0 && buf != 1) return fail(GF_EINVAL, "buffer must be 0 or 1");
    if (n < 0 || (n > 0 && (!patches || !tags))) return fail(GF_EINVAL, "invalid patch or tag list");
    if (!(eta_tol >= 0.0)) return fail(GF_EINVAL, "eta_tol must be >= 0");
    for (int k = 0; k < n; k++)
        if (patches[k] < 0 || patches[k] >= p->capacity || !(p->h_desc[patches[k]].dx > 0.0))
            return fail(GF_EINVAL, "entry %d: slot %d is not described", k, patches[k]);
    if (n == 0) return GF_OK;
    const size_t stride = ((sizeof(int) * (size_t)n + 255) / 256) * 256;
    if (int r = ensure_scratch(p, 2 * stride)) return r;
    char* d = (char*)p->tag_buf;
    HIPCHK(hipMemcpyAsync(d, patches, sizeof(int) * n, hipMemcpyHostToDevice, g.stream));
    gf::TagArgs a{};
    a.desc = p->desc;
    a.q = p->q[buf];
    a.aux = p->aux;
    a.patches = (const int*)d;
    a.n = n;
    a.out = (int*)(d + stride);
    a.drytol = g.P.dry_tolerance;
    HIPCHK(gf::launch_tag_eta_gradient(p->mx, p->my, a, eta_tol, g.stream));
    HIPCHK(hipMemcpyAsync(tags, a.out, sizeof(int) * n, hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    return GF_OK;
} GF_CATCH_STATUS

int gf_pool_copy(gf_pool* p, int n, const int* slots, int src_buf, int dst_buf)
try {
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    if ((src_buf != 0 && src_buf != 1) || (dst_buf != 0 && dst_buf != 1) || src_buf == dst_buf)
        return fail(GF_EINVAL, "buffers must be 0 and 1, in either order");
    if (n < 0 || (n > 0 && !slots)) return fail(GF_EINVAL, "invalid slot list");
    for (int k = 0; k < n; k++)
        if (slots[k] < 0 || slots[k] >= p->capacity) return fail(GF_EINVAL, "slot %d outside the pool", slots[k]);
    if (n == 0) return GF_OK;
    if (int r = ensure_scratch(p, sizeof(int) * (size_t)n)) return r;
    int* d = (int*)p->tag_buf;
    HIPCHK(hipMemcpyAsync(d, slots, sizeof(int) * n, hipMemcpyHostToDevice, g.stream));
    HIPCHK(gf::launch_copy_slots(p->q[src_buf], p->q[dst_buf], d, n, (int)p->np2, g.stream));
    // the destination buffer's depth maxima, and no steady-dry skip on the
    // strength of what the buffer held before
    HIPCHK(gf::launch_hmax_list(p->mx, p->my, p->q[dst_buf], p->hmax2 + (size_t)dst_buf * p->capacity, p->dry_age,
                                d, n, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    return GF_OK;
} GF_CATCH_STATUS

int gf_pool_timeinterp(gf_pool* p, int n, const int* src, int buf_last, int buf_curr, double alpha,
                       const int* dst, int dst_buf)
try {
    if (!p) return fail(GF_EINVAL, "pool is NULL");
    if ((buf_last != 0 && buf_last != 1) || (buf_curr != 0 && buf_curr != 1) || (dst_buf != 0 && dst_buf != 1))
        return fail(GF_EINVAL, "buffers must be 0 or 1");
    if (n < 0 || (n > 0 && (!src || !dst))) return fail(GF_EINVAL, "invalid slot lists");
    if (p->mx < 4 || p->my < 4) return fail(GF_EUNSUPPORTED, "patch too small");
    std::vector<char> used(p->capacity, 0);
    for (int k = 0; k < n; k++) {
        if (src[k] < 0 || src[k] >= p->capacity || dst[k] < 0 || dst[k] >= p->capacity)
            return fail(GF_EINVAL, "entry %d: slot outside the pool", k);
        if (used[dst[k]]) return fail(GF_EINVAL, "destination slot %d appears twice", dst[k]);
        used[dst[k]] = 1;
    }
    // a destination in a source buffer may not be another entry's source
    if (dst_buf == buf_last || dst_buf == buf_curr)
        for (int k = 0; k < n; k++)
            if (used[src[k]] && src[k] != dst[k]) return fail(GF_EINVAL, "slot %d is both a source and another entry's destination", src[k]);
    if (n == 0) return GF_OK;
    int *dsrc = nullptr, *ddst = nullptr;
    std::vector<int> dst4(4 * (size_t)n, 0);
    for (int k = 0; k < n; k++) dst4[k] = dst[k];
    if (int r = regrid_lists(p, n, src, dst4.data(), &dsrc, &ddst)) return r;
    gf::TimeInterpArgs a{};
    a.qlast = p->q[buf_last];
    a.qcurr = p->q[buf_curr];
    a.qdst = p->q[dst_buf];
    a.src = dsrc;
    a.dst = ddst;
    a.n = n;
    a.alpha = alpha;
    HIPCHK(gf::launch_timeinterp(p->mx, p->my, a, g.stream));
    HIPCHK(gf::launch_hmax_list(p->mx, p->my, p->q[dst_buf], p->hmax2 + (size_t)dst_buf * p->capacity, p->dry_age,
                                ddst, n, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    return GF_OK;
} GF_CATCH_STATUS

int gf_gauge_file_create(const char* dir, int id, double xc, double yc)
try {
    if (!dir) return fail(GF_EINVAL, "dir is NULL");
    char name[32];
    std::snprintf(name, sizeof name, "gauge%05d.txt", id);
    std::string s;
    gf::ascii::gauge_header(s, id, xc, yc);
    FILE* f = std::fopen(join_path(dir, name).c_str(), "wb");
    if (!f) return fail(GF_EIO, "cannot write %s", name);
    const bool ok = std::fwrite(s.data(), 1, s.size(), f) == s.size();
    if (std::fclose(f) != 0 || !ok) return fail(GF_EIO, "cannot write %s", name);
    return GF_OK;
} GF_CATCH_STATUS

int gf_gauge_file_append(const char* dir, int id, int n, const int* level, const double* t, const double* rec)
try {
    if (!dir) return fail(GF_EINVAL, "dir is NULL");
    if (n < 0 || (n > 0 && (!level || !t || !rec))) return fail(GF_EINVAL, "invalid records");
    char name[32];
    std::snprintf(name, sizeof name, "gauge%05d.txt", id);
    std::string s;
    for (int k = 0; k < n; k++)
        gf::ascii::gauge_line(s, level[k], t[k], rec[4 * k], rec[4 * k + 1], rec[4 * k + 2], rec[4 * k + 3]);
    FILE* f = std::fopen(join_path(dir, name).c_str(), "ab");
    if (!f) return fail(GF_EIO, "cannot write %s", name);
    const bool ok = std::fwrite(s.data(), 1, s.size(), f) == s.size();
    if (std::fclose(f) != 0 || !ok) return fail(GF_EIO, "cannot write %s", name);
    return GF_OK;
} GF_CATCH_STATUS

}  // extern "C"
