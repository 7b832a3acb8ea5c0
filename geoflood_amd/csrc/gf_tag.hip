// gf_tag.hip -- refinement tagging of resident patches on the device.
//
// Restates fc2d_geoclaw_flag2refine (fclaw2d_source/
// fc2d_geoflood_flag2refine.f90:7-130) per cell and its two patch-level
// callers: fc2d_geoclaw_fort_tag4refinement (a patch is tagged when any
// interior cell flags, fc2d_geoclaw_tag4refinement_fort.f90:1-59) and
// fc2d_geoclaw_fort_tag4coarsening (four siblings may coarsen when no cell of
// any of them flags under the coarsening thresholds,
// fc2d_geoclaw_tag4coarsening_fort.f90:1-110).  The reference scans cells in
// order and stops at the first flag; the result is an OR over cells, so the
// device evaluates every cell in parallel and ORs.  One workgroup per patch
// (or per family of four), reading h, hu, hv and b of the interior: an
// HBM-bound pass of 32 B per cell.
#include <hip/hip_runtime.h>

#include "gf_ghost.h"
#include "gf_launch.h"

namespace gf {

namespace {

// fc2d_geoclaw_flag2refine at one cell: 1 refine, 0 not.
__device__ int flag2refine(const TagCriteria& r, double drytol, double h, double hu, double hv, double b,
                           double x1, double y1, double x2, double y2, double t, int level, int maxlevel,
                           int init_flag, bool coarsening)
{
    if (init_flag != 0 && coarsening) return 0;  // :27-31
    for (int m = 0; m < r.mregions; m++) {       // forced refinement in a region (:42-50)
        const double* g = r.region + 8 * m;
        if (level < (int)g[6] && t >= g[4] && t <= g[5])
            if (x2 > g[0] && x1 < g[2] && y2 > g[1] && y1 < g[3]) return 1;
    }
    const int max_num_speeds = r.num_speeds < maxlevel ? r.num_speeds : maxlevel;
    const double momentum = sqrt(hu * hu + hv * hv);
    const double eta = h + b;
    const int th_factor = coarsening ? 0 : 1;  // an INTEGER in the reference: 0.5 -> 0 (:59-61)
    for (int m = 0; m < r.num_flowgrades; m++) {
        bool needed = false;
        const int var = r.fg_var[m];
        if (var == 1) needed = h > r.fg_value[m];
        else if (var == 2) needed = momentum > r.fg_value[m];
        else if (var == 3 && h > drytol) needed = fabs(eta) > r.fg_value[m];
        if (needed && level < r.fg_minlevel[m]) return 1;
    }
    double speed = 0.0;
    if (h > drytol) {
        speed = momentum / h;
        for (int n = 1; n <= max_num_speeds; n++)
            if (speed > th_factor * r.speed_tolerance[n - 1] && level <= n) return 1;
    }
    if (r.num_flowgrades == 0 && h > drytol) {
        if (fabs(eta - r.sea_level) > th_factor * r.wave_tolerance) {
            if (level < r.max_level_deep) return 1;
            if (fabs(b) < r.deep_depth) return 1;
        }
        if (fabs(b) > r.deep_depth && speed < r.speed_tolerance[0]) return 0;
        if (speed / h > r.max_velocity_depth_product) return 1;
    }
    return 0;
}

// Any interior cell of patch `patch` (buffer q) flags?
template <int MX, int MY>
__device__ bool patch_flags(const TagArgs& a, int patch, bool coarsening)
{
    constexpr int NXP = MX + 2 * MBC, NP2 = NXP * (MY + 2 * MBC);
    const PatchDesc d = a.desc[patch];
    const double* q = a.q + (size_t)patch * 3 * NP2;
    const double* b = a.aux + (size_t)patch * NP2;
    int any = 0;
    for (int c = threadIdx.x; c < MX * MY; c += blockDim.x) {
        const int i = c % MX + 1, j = c / MX + 1;
        const int k = (j + 1) * NXP + (i + 1);
        const double x1 = d.xlo + (i - 1) * d.dx, x2 = d.xlo + i * d.dx;
        const double y1 = d.ylo + (j - 1) * d.dy, y2 = d.ylo + j * d.dy;
        any |= flag2refine(*a.crit, a.drytol, q[k], q[NP2 + k], q[2 * NP2 + k], b[k], x1, y1, x2, y2, a.t,
                           a.level, a.maxlevel, a.init_flag, coarsening) > 0;
    }
    return __syncthreads_or(any) != 0;
}

template <int MX, int MY>
__global__ void __launch_bounds__(256) k_tag_refine(TagArgs a)
{
    const int k = blockIdx.x;
    if (k >= a.n) return;
    const bool f = patch_flags<MX, MY>(a, a.patches[k], false);
    if (threadIdx.x == 0) a.out[k] = f ? 1 : 0;
}

template <int MX, int MY>
__global__ void __launch_bounds__(256) k_tag_coarsen(TagArgs a)
{
    const int k = blockIdx.x;
    if (k >= a.n) return;
    bool f = false;
    for (int s = 0; s < 4; s++) f = f || patch_flags<MX, MY>(a, a.patches[4 * k + s], true);
    if (threadIdx.x == 0) a.out[k] = f ? 0 : 1;
}

// Surface-gradient criterion of an overland-flood run (no counterpart in the
// reference's flag2refine): interior cell (i, j) flags when it is wet
// (h > drytol) or when the centred differences of eta = h + b give
//   ((eta(i+1,j) - eta(i-1,j)) / (2 dx))^2 + ((eta(i,j+1) - eta(i,j-1)) / (2 dy))^2 > tol2,
// tol2 = |grad eta| threshold squared.  A neighbour beyond the interior has
// its h where the halo pull reads it (a neighbour's interior, a wall mirror,
// or the patch's own stored ghost on stored and inflow sides) and its b from
// the patch's own aux ghost cells, as the step reads them.
template <int MX, int MY>
__device__ __forceinline__ double eta_at(const TagArgs& a, const PatchDesc& d, int patch, int fi, int fj)
{
    using G = Geo<MX, MY>;
    const double* q = a.q + (size_t)patch * 3 * G::NP2;
    const double b = a.aux[(size_t)patch * G::NP2 + G::ti(fi, fj)];
    if (fi >= 1 && fi <= MX && fj >= 1 && fj <= MY) return q[G::ti(fi, fj)] + b;
    int sp, si, sj;
    bool nu, nv;
    if (ghost_source<MX, MY>(d.nbr, patch, fi, fj, sp, si, sj, nu, nv))
        return a.q[(size_t)sp * 3 * G::NP2 + G::ti(si, sj)] + b;
    return q[G::ti(fi, fj)] + b;
}

template <int MX, int MY>
__global__ void __launch_bounds__(256) k_tag_eta_grad(TagArgs a, double tol2)
{
    using G = Geo<MX, MY>;
    const int k = blockIdx.x;
    if (k >= a.n) return;
    const int patch = a.patches[k];
    const PatchDesc d = a.desc[patch];
    const double* q = a.q + (size_t)patch * 3 * G::NP2;
    int any = 0;
    for (int c = threadIdx.x; c < MX * MY; c += blockDim.x) {
        const int i = c % MX + 1, j = c / MX + 1;
        if (q[G::ti(i, j)] > a.drytol) {
            any = 1;
            continue;
        }
        const double gx = (eta_at<MX, MY>(a, d, patch, i + 1, j) - eta_at<MX, MY>(a, d, patch, i - 1, j)) / (2.0 * d.dx);
        const double gy = (eta_at<MX, MY>(a, d, patch, i, j + 1) - eta_at<MX, MY>(a, d, patch, i, j - 1)) / (2.0 * d.dy);
        if (gx * gx + gy * gy > tol2) any = 1;
    }
    any = __syncthreads_or(any);
    if (threadIdx.x == 0) a.out[k] = any ? 1 : 0;
}

}  // namespace

hipError_t launch_tag_eta_gradient(int mx, int my, const TagArgs& a, double eta_tol, hipStream_t st)
{
    if (a.n <= 0) return hipSuccess;
    if (mx != my) return hipErrorInvalidValue;
    const double tol2 = eta_tol * eta_tol;
    switch (mx) {
    case 8: k_tag_eta_grad<8, 8><<<a.n, 256, 0, st>>>(a, tol2); break;
    case 16: k_tag_eta_grad<16, 16><<<a.n, 256, 0, st>>>(a, tol2); break;
    case 32: k_tag_eta_grad<32, 32><<<a.n, 256, 0, st>>>(a, tol2); break;
    case 64: k_tag_eta_grad<64, 64><<<a.n, 256, 0, st>>>(a, tol2); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_tag(int mx, int my, bool coarsen, const TagArgs& a, hipStream_t st)
{
    if (a.n <= 0) return hipSuccess;
    if (mx != my) return hipErrorInvalidValue;
    switch (mx) {
    case 8: coarsen ? k_tag_coarsen<8, 8><<<a.n, 256, 0, st>>>(a) : k_tag_refine<8, 8><<<a.n, 256, 0, st>>>(a); break;
    case 16: coarsen ? k_tag_coarsen<16, 16><<<a.n, 256, 0, st>>>(a) : k_tag_refine<16, 16><<<a.n, 256, 0, st>>>(a); break;
    case 32: coarsen ? k_tag_coarsen<32, 32><<<a.n, 256, 0, st>>>(a) : k_tag_refine<32, 32><<<a.n, 256, 0, st>>>(a); break;
    case 64: coarsen ? k_tag_coarsen<64, 64><<<a.n, 256, 0, st>>>(a) : k_tag_refine<64, 64><<<a.n, 256, 0, st>>>(a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace gf
