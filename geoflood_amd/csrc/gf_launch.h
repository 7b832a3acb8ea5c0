// gf_launch.h -- kernel argument blocks and host-side launchers of the
// patch-step kernels (defined in gf_step.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "gf_kernels.h"

namespace gf {

struct StepArgs {
    const PatchDesc* desc;
    const int* list;      // optional work list of patch indices
    int first;            // first patch when list == nullptr
    int count;            // number of patches to process
    double* q0;           // q buffer 0
    double* q1;           // q buffer 1 (nullptr: update in place)
    const double* aux;
    double* cfl_out;      // per patch, MAXCFL_CAP applied
    double* hmax_out;     // per patch max interior depth after the step (nullable)
    ScenCtrl* ctrl;       // per-scenario dt / buffer index (nullable)
    double dt;            // when ctrl == nullptr
    int cur;              // when ctrl == nullptr: read buffer cur, write buffer cur^1
    Params P;
    unsigned long long* tile_cfl;  // per patch: tile fold slots (tiled variants)
    unsigned long long* tile_h;
    int* tile_cnt;
};

struct GhostArgs {
    const PatchDesc* desc;
    int first, count;
    double* q0;
    double* q1;
    ScenCtrl* ctrl;  // when set, each patch uses its scenario's current buffer
    int cur;         // otherwise this buffer
};

// Gauge samples: gauge g at (xy[2g], xy[2g+1]) inside patch slot[g] of
// buffer q; out[4g..4g+3] = h, hu, hv, b there.
struct GaugeArgs {
    const PatchDesc* desc;
    const double* q;    // the state buffer sampled
    const double* aux;
    const int* slot;
    const double* xy;
    double* out;
    int n;
    double dry_tolerance;
};

// Regrid between levels on resident patches (refinement ratio 2): family f
// has the coarse patch in slot cs[f] and its children in slots fs[4f + g],
// g = igrid = 0..3 (x fastest), all in buffer q.
//
// Coarse/fine ghost transfers also run in a resident mode (ctrl != nullptr),
// as nodes of the attempt graph: link l's buffer is the current one of the
// coarse patch's scenario (q, or q1 when its ``cur`` is 1; links of inactive
// scenarios are skipped), and the links are [range[0], range[1]) of the
// arrays, read from device memory, taken grid-stride by n workgroups.
struct RegridArgs {
    const PatchDesc* desc;
    double* q;
    const double* aux;
    const int* cs;
    const int* fs;
    const int* side;   // coarse/fine ghost transfers: 0..3 face, 4..7 corner
    const int* igrid;  // and the fine neighbour's half along a face
    int n;
    double dry_tolerance;
    double* q1;             // resident mode: the other q buffer
    const ScenCtrl* ctrl;   // resident mode: per-scenario current buffer
    const int* range;       // nullable: links [range[0], range[1]) instead of [0, n)
};

// Time interpolation of the two interior border layers: slot dst[k] of qdst
// from slot src[k] of qlast and qcurr.
struct TimeInterpArgs {
    const double* qlast;
    const double* qcurr;
    double* qdst;
    const int* src;
    const int* dst;
    int n;
    double alpha;
};

// Hydrograph inflow ghost cells of the patches in list[0 .. *list_n) that lie
// in [first, last): each patch's ghost ring is rebuilt in its q buffer (and
// aux) by ForestClaw's sequence -- same-level copies, then bc2 sides left,
// right, bottom, top.  buf < 0: each patch uses its scenario's current buffer
// and time and skips inactive scenarios; otherwise buffer buf at time t.
struct InflowArgs {
    const PatchDesc* desc;
    const ScenCtrl* ctrl;
    double* q0;
    double* q1;
    double* aux;
    const int* list;
    const int* list_n;
    int first, last;
    int buf;
    double t;
    Params P;
    const int* range;  // nullable: entries [range[0], range[1]) of list (one level) instead of [0, *list_n)
};

// Dry-list entries carry the patch's current buffer: patch | cur << 30.
constexpr int DRY_CUR_SHIFT = 30;

// Device counters of the batched path (BatchArgs::counters).
enum { CNT_WET = 0, CNT_DRY = 1, NEXT_WET = 2, CNT_SKIP = 3, NCOUNTERS = 4 };

// Steady dry patches.  A dry patch whose cells are all below dry_tolerance
// steps to (max(h,0), 0, 0), which is idempotent: after two such steps into
// alternating buffers both buffers hold the same interior, and every later
// dry step leaves it unchanged.  dry_age[p] counts those steps (saturating at
// DRY_STEADY) in its low bits and remembers the buffer last written in bit
// DRY_LAST_BIT (a retake writes the same buffer again and does not count);
// the classification skips patches at DRY_STEADY.  Any other write of a
// patch's state resets its entry to 0.
constexpr int DRY_STEADY = 2;
constexpr int DRY_LAST_BIT = 4;

// Arguments of the batched (work-list) path: classify -> wet || dry -> control.
struct BatchArgs {
    const PatchDesc* desc;
    int n;                        // patches in the pool
    int capacity;
    double* q0;
    double* q1;
    const double* aux;
    double* hmax;                 // [2][capacity] max interior depth of each buffer's state
    unsigned long long* cflbits;  // per scenario: max CFL of the attempt (bit pattern, CFL >= 0)
    ScenCtrl* ctrl;
    int nscen;
    int* list_wet;
    int* list_dry;
    int* counters;                // NCOUNTERS ints
    long long* stats;             // accumulated active, dry (steady included), attempts, steady
    int* dry_age;                 // per patch, see DRY_STEADY
    Params P;
    unsigned long long* tile_cfl; // per patch: tile fold slots (tiled variants)
    unsigned long long* tile_h;
    int* tile_cnt;
};

// Refinement criteria in device memory (the arrays follow the struct in
// the same allocation).
struct TagCriteria {
    double wave_tolerance, deep_depth, max_velocity_depth_product, sea_level;
    int max_level_deep, num_speeds, num_flowgrades, mregions;
    const double* speed_tolerance;
    const double* fg_value;
    const int* fg_var;
    const int* fg_minlevel;
    const double* region;  // 8 per region: xlow, ylow, xhi, yhi, tlow, thi, minlevel, maxlevel
};

struct TagArgs {
    const PatchDesc* desc;
    const double* q;        // the q buffer to examine
    const double* aux;
    const TagCriteria* crit;
    const int* patches;     // n patches (refine) or 4n (coarsen: four siblings per family)
    int n;
    int* out;
    double t, drytol;
    int level, maxlevel, init_flag;
};
hipError_t launch_tag(int mx, int my, bool coarsen, const TagArgs& a, hipStream_t st);
// Surface-gradient tagging: out[k] = 1 when an interior cell of patch
// patches[k] (buffer q) has h > drytol or |grad eta| > eta_tol (centred
// differences of eta = h + b, neighbours read as the halo pull reads them).
hipError_t launch_tag_eta_gradient(int mx, int my, const TagArgs& a, double eta_tol, hipStream_t st);

// A step variant: patch shape, tile shape, threads per patch (tile),
// patches per workgroup of the persistent kernel, and where the x-face flux
// increments wait for the y phase (1 registers, 0 LDS).
struct StepVariant {
    int mx, my, tx, ty, nt, npp, freg;
};
int num_variants(int mx, int my);
bool get_variant(int mx, int my, int k, StepVariant* out);
// Resident workgroups per CU of the persistent step kernel of a variant.
int wet_occupancy(int mx, int my, int variant);

bool supported_shape(int mx, int my);
// Work lists of an attempt over slots [0, da->n), read from the device
// arguments (a graph replay sees the current count); grid for `slots` slots.
hipError_t launch_classify(const BatchArgs* da, int slots, hipStream_t st);
// da: the BatchArgs in device memory (read by the persistent kernel on every iteration)
// def: the method is the default one (order 2, transverse 2, MC limiter);
// its kernels have the method fixed at compile time.
bool default_method(const Params& P);
// Active tiles then dry patches from one persistent kernel (see k_all).
hipError_t launch_all(int mx, int my, int variant, bool def, const BatchArgs* da, int grid, hipStream_t st);
// The dt controller of scenarios [0, da->nscen) (device arguments); grid for
// `scenarios` scenario slots.
hipError_t launch_batch_control(const BatchArgs* da, int scenarios, hipStream_t st);
hipError_t launch_average2coarse(int mx, int my, const RegridArgs& a, hipStream_t st);
hipError_t launch_interpolate2fine(int mx, int my, const RegridArgs& a, hipStream_t st);
hipError_t launch_copy_slots(const double* src, double* dst, const int* slots, int n, int np2, hipStream_t st);
hipError_t launch_average_ghost(int mx, int my, const RegridArgs& a, hipStream_t st);
hipError_t launch_interpolate_ghost(int mx, int my, const RegridArgs& a, hipStream_t st);
hipError_t launch_timeinterp(int mx, int my, const TimeInterpArgs& a, hipStream_t st);
// hmax[slot] of buffer q for the listed slots; clears their dry_age entries.
hipError_t launch_hmax_list(int mx, int my, const double* q, double* hmax, int* dry_age, const int* slots, int n,
                            hipStream_t st);
hipError_t launch_frame_pack(int mx, int my, double* q, const double* aux, int first, int n, double* out,
                             hipStream_t st);
hipError_t launch_gauges(int mx, int my, const GaugeArgs& a, hipStream_t st);
hipError_t launch_hmax(int mx, int my, const double* q, double* hmax, int first, int n, hipStream_t st);
hipError_t launch_step(int mx, int my, int variant, const StepArgs& a, hipStream_t st);
hipError_t launch_ghost(int mx, int my, const GhostArgs& a, hipStream_t st);
hipError_t launch_control(ScenCtrl* ctrl, int nscen, const double* cfl, hipStream_t st);
hipError_t launch_inflow(int mx, int my, const InflowArgs& a, int grid, hipStream_t st);
// q2 (nullable): a second q buffer that receives the same state
hipError_t launch_pack(const double* stage_q, const double* stage_aux, double* q, double* q2, double* aux,
                       int first, int n, int np2, int maux, hipStream_t st);
hipError_t launch_unpack(double* stage_q, const double* q, int first, int n, int np2, hipStream_t st);
// descriptors of n host patches: dx[k], dy[k], every side stored
hipError_t launch_desc_fill(PatchDesc* d, const double* dx, const double* dy, int n, hipStream_t st);
// the step's powers (gf_pow.h) of n device values: which 0 -> x^(7/3), 1 -> x^(2/3)
hipError_t launch_eval_powers(int which, const double* x, double* out, int n, hipStream_t st);
// n 4-byte words set to zero (a kernel, so that a captured stream holds no memset node)
hipError_t launch_zero_words(int* p, size_t n, hipStream_t st);
// q (one buffer) and aux of slot src[k] to slot dst[k]
hipError_t launch_move_slots(double* q, double* aux, const int* src, const int* dst, int n, int np2, hipStream_t st);
// desc[first .. first + n).scen = scen
hipError_t launch_set_scen(PatchDesc* desc, int first, int n, int scen, hipStream_t st);
hipError_t launch_max(const double* v, int n, double* out, hipStream_t st);

}  // namespace gf
