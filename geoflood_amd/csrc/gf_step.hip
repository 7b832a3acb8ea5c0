// gf_step.hip -- fused FP64 patch step for gfx950 (MI355X).
//
// One workgroup advances one patch by one step:
//   tile load (+ b4step2)  ->  x-faces: rpn2, MC limiter, flux corrections
//   -> x-transverse rpt2   ->  y-faces: rpn2, limiter, corrections
//   -> y-transverse rpt2   ->  conservative update + src2 (Manning)
// together with the patch's CFL number (max |s| dt/dx over the reference's
// face set, including the faces in the ghost rows/columns).
//
// Work decomposition.  The patch has (mx+3) x (my+3) "positions" (i, j),
// Fortran indices i, j in [0, mx+2] x [0, my+2].  Position (i, j) owns x-face
// (i, j) (between cells (i-1, j) and (i, j)), y-face (i, j) (between (i, j-1)
// and (i, j)) and cell (i, j).  Thread t owns positions t, t+NT, ...; all
// per-face and per-cell intermediates stay in the owning thread's registers,
// and only what a neighbouring position needs crosses through LDS:
//   unlimited f-waves (limiter), amdq* (transverse split of the face to the
//   right/top), the G/F transverse increments of the cell below/left, and the
//   final fm/gm of the face to the right/top.
// Every floating-point value is formed with the same operations in the same
// order as the reference Fortran (geoclaw_step2_fort.f90, geoclaw_flux2fw_
// fort.f, rpn2_geoclaw.f, rpt2_geoclaw.f, geoclaw_src2_fort.f90), so the
// results equal the reference's up to the sign of exact zeros and the rare
// calls in which the host libm's pow misrounds (src2's power is correctly
// rounded here, gf_pow.h).
#include <hip/hip_runtime.h>

#include <algorithm>

#include <cstddef>

#include "gf_ghost.h"
#include "gf_inflow.h"
#include "gf_launch.h"

namespace gf {


// Maximum in which a NaN operand wins (fmax would drop it), so that a
// non-finite state reaches the scenario's controller.  +0 wins over -0.
__device__ __forceinline__ double max_nan(double a, double b) { return (b > a || b != b) ? b : a; }

__device__ __forceinline__ double quiet_nan() { return __longlong_as_double(0x7ff8000000000000ll); }

// The three values are finite (no NaN, no Inf).
__device__ __forceinline__ bool finite3(double a, double b, double c)
{
    return (a - a == 0.0) && (b - b == 0.0) && (c - c == 0.0);
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = max_nan(v, o);
    }
    return v;
}

// Compressed f-waves of one face: wave 1 (3), normal momentum of wave 2 (1),
// wave 3 (3).  The other two components of wave 2 are exact zeros
// (geoclaw_riemann_utils.f:289-299: fw(1,2) = beta(2)*r(2,2) = 0, fw(3,2) = 0).
__device__ __forceinline__ void compress(const double w[3][3], double c[7])
{
    c[0] = w[0][0]; c[1] = w[0][1]; c[2] = w[0][2];
    c[3] = w[1][1];
    c[4] = w[2][0]; c[5] = w[2][1]; c[6] = w[2][2];
}

// Dot product of two 3-vectors given in direction-local order (h, n, t),
// summed in the reference's physical component order (h, hu, hv):
// x-direction (h, n, t) = (h, hu, hv); y-direction (h, n, t) = (h, hv, hu).
template <int DIR>
__device__ __forceinline__ double pdot(double a0, double a1, double a2, double b0, double b1, double b2)
{
    if (DIR == 0) return a0 * b0 + a1 * b1 + a2 * b2;
    return a0 * b0 + a2 * b2 + a1 * b1;
}

// Limiter, second-order correction and fluctuation split of one face
// (geoclaw_flux2fw_fort.f:143-224 with geoclaw_limiter_fort.f:34-58 and
// philim.f).  wc: own compressed unlimited waves; wl / wr: those of the faces
// before / after along the sweep.  Outputs in direction-local order: faddm,
// faddp (flux increments) and amdq*, apdq* (fluctuations handed to rpt2).
// DEF: the default method (order 2, transverse 2, MC limiter on all three
// families) fixed at compile time; otherwise the method is read from P.
template <int DIR, bool DEF>
__device__ __forceinline__ void face_second_order(const double wc[7], const double s[3],
                                                  const double wl[7], const double wr[7],
                                                  double dtdx, const Params& P,
                                                  double faddm[3], double faddp[3],
                                                  double amq[3], double apq[3])
{
    double w[3][3] = {{wc[0], wc[1], wc[2]}, {0.0, wc[3], 0.0}, {wc[4], wc[5], wc[6]}};
    double amdq[3], apdq[3];
    fluctuations(s, w, amdq, apdq);
    const int order = DEF ? 2 : P.order;
    const int lim0 = DEF ? 4 : P.mthlim[0], lim1 = DEF ? 4 : P.mthlim[1], lim2 = DEF ? 4 : P.mthlim[2];
    if (order == 1) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            faddm[k] = amdq[k];
            faddp[k] = -apdq[k];
            amq[k] = amdq[k];
            apq[k] = apdq[k];
        }
        return;
    }
    // limiter on the three wave families with the unlimited neighbours
    {
        // wave 1
        if (lim0 != 0) {
            const double wn2 = pdot<DIR>(wc[0], wc[1], wc[2], wc[0], wc[1], wc[2]);
            if (wn2 != 0.0) {
                const double dotl = pdot<DIR>(wl[0], wl[1], wl[2], wc[0], wc[1], wc[2]);
                const double dotr = pdot<DIR>(wc[0], wc[1], wc[2], wr[0], wr[1], wr[2]);
                const double lim = philim(wn2, s[0] > 0.0 ? dotl : dotr, lim0);
                w[0][0] = lim * w[0][0]; w[0][1] = lim * w[0][1]; w[0][2] = lim * w[0][2];
            }
        }
        // wave 2 (only its normal-momentum component is non-zero)
        if (lim1 != 0) {
            const double wn2 = wc[3] * wc[3];
            if (wn2 != 0.0) {
                const double dotl = wl[3] * wc[3];
                const double dotr = wc[3] * wr[3];
                const double lim = philim(wn2, s[1] > 0.0 ? dotl : dotr, lim1);
                w[1][1] = lim * w[1][1];
            }
        }
        // wave 3
        if (lim2 != 0) {
            const double wn2 = pdot<DIR>(wc[4], wc[5], wc[6], wc[4], wc[5], wc[6]);
            if (wn2 != 0.0) {
                const double dotl = pdot<DIR>(wl[4], wl[5], wl[6], wc[4], wc[5], wc[6]);
                const double dotr = pdot<DIR>(wc[4], wc[5], wc[6], wr[4], wr[5], wr[6]);
                const double lim = philim(wn2, s[2] > 0.0 ? dotl : dotr, lim2);
                w[2][0] = lim * w[2][0]; w[2][1] = lim * w[2][1]; w[2][2] = lim * w[2][2];
            }
        }
    }
    double c[3];
#pragma unroll
    for (int mw = 0; mw < 3; mw++) c[mw] = copysign(1.0, s[mw]) * (1.0 - fabs(s[mw]) * dtdx);
    double cq[3];
    cq[0] = c[0] * w[0][0] + c[2] * w[2][0];
    cq[1] = c[0] * w[0][1] + c[1] * w[1][1] + c[2] * w[2][1];
    cq[2] = c[0] * w[0][2] + c[2] * w[2][2];
    const bool fold = DEF ? true : (P.transverse == 2);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        faddm[k] = amdq[k] + 0.5 * cq[k];
        faddp[k] = -apdq[k] + 0.5 * cq[k];
        amq[k] = fold ? amdq[k] + cq[k] : amdq[k];
        apq[k] = fold ? apdq[k] - cq[k] : apdq[k];
    }
}

// Transverse increments of one cell from the two normal fluctuations that
// its state splits (geoclaw_flux2fw_fort.f:232-265): G1 (below/left) and G2
// (above/right), direction-local order, hd = dt/dx / 2.
__device__ __forceinline__ void cell_transverse(const Rpt2Cell& C, const double am_next[3],
                                                const double ap_own[3], double hd,
                                                double G1[3], double G2[3])
{
    double bm1[3], bp1[3], bm2[3], bp2[3];
    rpt2_apply(C, am_next, bm1, bp1);
    rpt2_apply(C, ap_own, bm2, bp2);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        G1[k] = -(hd * bm1[k]) - hd * bm2[k];
        G2[k] = -(hd * bp1[k]) - hd * bp2[k];
    }
}

// Maximum over a group of NT threads (the workgroup, or one of the groups
// of a workgroup that steps several patches; ltid is the thread's index in
// its group, NT a multiple of 64) with one barrier of the whole workgroup;
// the result is valid in the group's thread 0.
template <int NT>
__device__ __forceinline__ double block_max(double v, double* red, int ltid)
{
    constexpr int NWAVE = NT / 64;
    v = wave_max(v);
    __syncthreads();
    if ((ltid & 63) == 0) red[ltid >> 6] = v;
    __syncthreads();
    double r = red[0];
#pragma unroll
    for (int k = 1; k < NWAVE; k++) r = max_nan(r, red[k]);
    return r;
}

// Group maxima of two values with one barrier (results valid in every
// thread of the group).  red holds 2 * NT / 64 doubles per group; the caller
// guarantees that no thread still reads red from an earlier use (a barrier
// separates them).
template <int NT>
__device__ __forceinline__ void block_max2(double& a, double& b, double* red, int ltid)
{
    constexpr int NWAVE = NT / 64;
    a = wave_max(a);
    b = wave_max(b);
    if ((ltid & 63) == 0) {
        red[ltid >> 6] = a;
        red[NWAVE + (ltid >> 6)] = b;
    }
    __syncthreads();
    a = red[0];
    b = red[NWAVE];
#pragma unroll
    for (int k = 1; k < NWAVE; k++) {
        a = max_nan(a, red[k]);
        b = max_nan(b, red[NWAVE + k]);
    }
}

// The whole step of one tile of a patch by one workgroup of NT threads.  A
// tile is the TX x TY block of interior cells at offset (ox, oy) of an
// MX x MY patch (TX = MX, TY = MY: the whole patch) with its own 2-cell
// halo: the update of an interior cell reaches two cells in each direction
// (limiter and transverse corrections), so a tile computes exactly what the
// whole-patch step computes for its cells.  q_src / q_dst are the pool's
// source and destination buffers (q_dst == q_src updates in place).  With
// PULL the halo cells outside the patch are gathered from the neighbours'
// interiors in q_src (same-level exchange and physical sides fused into the
// load); otherwise the stored ghost cells are used; halo cells inside the
// patch are its own interior.  Thread 0 receives the tile's maximum CFL
// number (uncapped) and the maximum interior depth after the step.
//
// A workgroup may step several patches at once, one per group of NT
// threads (ltid: the thread's index in its group; T, E, F, red: the group's
// slices of LDS).  The groups pass every barrier together, so the all-dry
// fast path is taken only when every group's tile is dry (the full path
// gives the same values on a dry tile).  A group without a patch (valid
// false: the last item of the work list) computes on a zero tile and writes
// nothing.  FREG: the x-face flux increments stay in the owning thread's
// registers between the x and the y phase (the thread that forms a face's
// increments is the one that reads them back) instead of in LDS -- the "LDS
// split" of a step variant.
//
// Hook: called by thread NT-1 (of group 0) right after the tile is in LDS:
// that thread has no position in the last slot, so its wave has slack in
// the phases that follow.
struct NoHook {
    __device__ void operator()() const {}
};

template <int MX, int MY, int TX, int TY, int NT, bool PULL, bool DEF, bool FREG = false, typename Hook = NoHook>
__device__ __forceinline__ void step_patch(double* T, double* E, double* F, double* red, const PatchDesc* dptr, int patch,
                           int ox, int oy, const double* q_src, double* q_dst, const double* aux,
                           double dt, const Params& P, double& cfl_res, double& hmax_res,
                           int ltid, bool valid, const Hook& hook = Hook())
{
    using G = Geo<TX, TY>;   // tile geometry (LDS)
    using GP = Geo<MX, MY>;  // patch geometry (HBM)
    constexpr int NXP = G::NXP, NP2 = G::NP2, PX = G::PX, NPOS = G::NPOS;
    constexpr int PNP2 = GP::NP2;
    constexpr int SLOTS = (NPOS + NT - 1) / NT;
    // Opaque copy of the thread index: inside the persistent loop this stops
    // the compiler from hoisting every index computation of the body out of
    // the loop, which would keep them live in registers across iterations.
    int tid = ltid;
    asm volatile("" : "+v"(tid));
    const double* src = q_src + (size_t)patch * 3 * PNP2;
    double* dst = q_dst + (size_t)patch * 3 * PNP2;
    const double* auxp = aux + (size_t)patch * PNP2;
    const double g = P.grav, drytol = P.dry_tolerance;
    const double dtdx = dt / dptr->dx;
    const double dtdy = dt / dptr->dy;

    // ---- tile load (+ halo gather) + b4step2 (geoclaw_b4step2_fort.f90:153-156) ----
    // Addresses of every cell this thread loads first, then all loads, then
    // the LDS stores: one memory round trip per tile instead of one per cell
    // group.
    constexpr int LS = (NP2 + NT - 1) / NT;
    const double* lbase[LS];
    int lk[LS], lpc[LS];
    bool lok[LS], lnu[LS], lnv[LS];
#pragma unroll
    for (int r = 0; r < LS; r++) {
        const int c = tid + r * NT;
        lok[r] = valid && c < NP2;
        const int cc = lok[r] ? c : 0;
        const int fi = ox + cc % NXP - 1, fj = oy + cc / NXP - 1;  // patch Fortran indices
        const int pc = GP::ti(fi, fj);
        const double* base = src;
        int k = pc;
        bool nu = false, nv = false;
        if (PULL) {
            const bool interior = (fi >= 1 && fi <= MX && fj >= 1 && fj <= MY);
            int sp = patch, si = fi, sj = fj;
            if (!interior && ghost_source<MX, MY>(dptr->nbr, patch, fi, fj, sp, si, sj, nu, nv)) {
                base = q_src + (size_t)sp * 3 * PNP2;
                k = GP::ti(si, sj);
            }
        }
        lbase[r] = base;
        lk[r] = k;
        lpc[r] = pc;
        lnu[r] = nu;
        lnv[r] = nv;
    }
    double lh[LS], lhu[LS], lhv[LS], lb[LS];
#pragma unroll
    for (int r = 0; r < LS; r++) {
        if (lok[r]) {
            lh[r] = lbase[r][lk[r]];
            lhu[r] = lbase[r][PNP2 + lk[r]];
            lhv[r] = lbase[r][2 * PNP2 + lk[r]];
            lb[r] = auxp[lpc[r]];
        }
    }
    int wet = 0;
#pragma unroll
    for (int r = 0; r < LS; r++) {
        if (lok[r]) {
            const int c = tid + r * NT;
            double h = lh[r], hu = lnu[r] ? -lhu[r] : lhu[r], hv = lnv[r] ? -lhv[r] : lhv[r];
            if (h < drytol) {
                h = fmaxf64(h, 0.0);
                hu = 0.0;
                hv = 0.0;
            }
            T[c] = h;
            T[NP2 + c] = hu;
            T[2 * NP2 + c] = hv;
            T[3 * NP2 + c] = lb[r];
            wet |= (h > drytol);
        } else if (!valid && tid + r * NT < NP2) {
            const int c = tid + r * NT;  // a group without a patch: a dry zero tile
            T[c] = 0.0;
            T[NP2 + c] = 0.0;
            T[2 * NP2 + c] = 0.0;
            T[3 * NP2 + c] = 0.0;
        }
    }
    const int anywet = __syncthreads_or(wet);
    static_assert(NT - 1 + NT * (SLOTS - 1) >= G::NPOS || SLOTS == 1,
                  "the hook thread must have no position in the last slot");
    if (threadIdx.x == NT - 1) hook();

    if (!anywet) {
        // Every face of the tile is dry on both sides: all fluctuations, waves
        // and transverse splits are zero and the update leaves q unchanged, so
        // the step reduces to b4step2 + src2 on the interior and CFL = 0.
        double hm = 0.0;
        bool bad = false;
        for (int c = tid; c < TX * TY; c += NT) {
            const int i = c % TX + 1, j = c / TX + 1;
            const int k = G::ti(i, j);
            const int pk = GP::ti(ox + i, oy + j);
            double h = T[k], hu = T[NP2 + k], hv = T[2 * NP2 + k];
            if (P.src_term > 0) src2_cell(h, hu, hv, T[3 * NP2 + k], dt, P);
            if (valid) {
                dst[pk] = h;
                dst[PNP2 + pk] = hu;
                dst[2 * PNP2 + pk] = hv;
            }
            hm = fmaxf64(hm, h);
            bad |= !finite3(h, hu, hv);
        }
        // a non-finite cell (a NaN depth is neither wet nor dry) makes the
        // tile's CFL number NaN, which stops the scenario
        if (bad) hm = quiet_nan();
        hm = block_max<NT>(hm, red, tid);
        if (tid == 0) {
            cfl_res = hm != hm ? quiet_nan() : 0.0;
            hmax_res = hm;
        }
        return;
    }

    int pi[SLOTS], pj[SLOTS];
    bool in[SLOTS];
#pragma unroll
    for (int s = 0; s < SLOTS; s++) {
        const int p = tid + s * NT;
        in[s] = p < NPOS;
        pi[s] = p % PX;
        pj[s] = p / PX;
    }

    double cfl = 0.0;
    // A face's own compressed waves go to LDS for its neighbours' limiter and
    // are read back from there by the face itself, so that they do not stay
    // in registers while the other slots solve their Riemann problems.
    double Sr[SLOTS][3];
    // each position's own cell (the right-hand cell of its x- and y-face and the
    // cell of both transverse splits): hu/h, hv/h and sqrt(g*h) once, for a wet
    // cell, instead of in each of those four places.  Kept for the large tiles
    // of up to 768 threads only: there one workgroup per CU is set by LDS and
    // the extra registers are free (up to 168 at three waves per SIMD), while
    // the other variants' occupancy is set by registers.
    constexpr bool OWN = TX >= 32 && TY >= 32 && NT <= 768;
    double CU[SLOTS], CV[SLOTS], CS[SLOTS];
#pragma unroll
    for (int s = 0; s < SLOTS; s++) {
        CU[s] = 0.0; CV[s] = 0.0; CS[s] = 0.0;
        if (OWN) {
            const int k = G::ti(pi[s], pj[s]);
            const double h = in[s] ? T[k] : 0.0;
            if (h > drytol) {
                CU[s] = T[NP2 + k] / h;
                CV[s] = T[2 * NP2 + k] / h;
                CS[s] = sqrt(g * h);
            }
        }
    }
    double AP[SLOTS][3], AM[SLOTS][3], GA[SLOTS][3], GB[SLOTS][3];
    double FX[FREG ? SLOTS : 1][6];  // FREG: the position's x-face increments (faddm, faddp)

    // ================= x-faces: normal Riemann problems =================
#pragma unroll
    for (int s = 0; s < SLOTS; s++) {
        const int i = pi[s], j = pj[s];
        double sw[3] = {0.0, 0.0, 0.0};
        double w[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
        if (in[s] && j <= TY + 1) {
            const int kL = G::ti(i - 1, j), kR = G::ti(i, j);
            rpn2_face<OWN>(T[kL], T[NP2 + kL], T[2 * NP2 + kL], T[3 * NP2 + kL],
                           T[kR], T[NP2 + kR], T[2 * NP2 + kR], T[3 * NP2 + kR], P, sw, w,
                           RightCell{CU[s], CV[s], CS[s]});
#pragma unroll
            for (int mw = 0; mw < 3; mw++) {
                const double v = dtdx * sw[mw];
                cfl = fmaxf64(cfl, v);
                cfl = fmaxf64(cfl, -v);
            }
        }
        double wc[7];
        compress(w, wc);
        Sr[s][0] = sw[0]; Sr[s][1] = sw[1]; Sr[s][2] = sw[2];
        if (in[s]) {
            const int p = tid + s * NT;
#pragma unroll
            for (int k = 0; k < 7; k++) E[k * NPOS + p] = wc[k];
        }
    }
    __syncthreads();

    // ================= x-faces: limiter + corrections =================
#pragma unroll
    for (int s = 0; s < SLOTS; s++) {
        const int i = pi[s], j = pj[s];
        const int p = tid + s * NT;
        if (in[s] && i >= 1 && i <= TX + 1 && j <= TY + 1) {
            double wl[7], wc[7], wr[7];
#pragma unroll
            for (int k = 0; k < 7; k++) {
                wl[k] = E[k * NPOS + p - 1];
                wc[k] = E[k * NPOS + p];
                wr[k] = E[k * NPOS + p + 1];
            }
            double fxm[3], fxp[3];
            face_second_order<0, DEF>(wc, Sr[s], wl, wr, dtdx, P, fxm, fxp, AM[s], AP[s]);
            if constexpr (FREG) {
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    FX[s][k] = fxm[k];
                    FX[s][3 + k] = fxp[k];
                }
            } else if (j >= 1 && j <= TY) {
                const int f = G::fi_(i, j);
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    F[k * G::NF + f] = fxm[k];
                    F[(3 + k) * G::NF + f] = fxp[k];
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < 3; k++) { AM[s][k] = 0.0; AP[s][k] = 0.0; }
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < SLOTS; s++)
        if (in[s]) {
            const int p = tid + s * NT;
#pragma unroll
            for (int k = 0; k < 3; k++) E[k * NPOS + p] = AM[s][k];
        }
    __syncthreads();

    // ================= x-transverse (cell-centred rpt2) =================
    const bool transverse = DEF ? true : P.transverse != 0;
    const double hdx = 0.5 * dtdx;
#pragma unroll
    for (int s = 0; s < SLOTS; s++) {
        const int i = pi[s], j = pj[s];
        const int p = tid + s * NT;
#pragma unroll
        for (int k = 0; k < 3; k++) { GA[s][k] = 0.0; GB[s][k] = 0.0; }
        if (transverse && in[s] && i >= 1 && i <= TX && j <= TY + 1) {
            const int kc = G::ti(i, j);
            const Rpt2Cell C =
                OWN ? rpt2_setup_c(T[kc], CU[s], CV[s], CS[s], T[3 * NP2 + kc], T[3 * NP2 + G::ti(i, j - 1)],
                                   T[3 * NP2 + G::ti(i, j + 1)], drytol)
                    : rpt2_setup(T[kc], T[NP2 + kc], T[2 * NP2 + kc], T[3 * NP2 + kc], T[3 * NP2 + G::ti(i, j - 1)],
                                 T[3 * NP2 + G::ti(i, j + 1)], g, drytol);
            const double amn[3] = {E[p + 1], E[NPOS + p + 1], E[2 * NPOS + p + 1]};
            cell_transverse(C, amn, AP[s], hdx, GA[s], GB[s]);  // GA = G1 (below), GB = G2 (above)
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < SLOTS; s++)
        if (in[s]) {
            const int p = tid + s * NT;
#pragma unroll
            for (int k = 0; k < 3; k++) E[k * NPOS + p] = GB[s][k];
        }
    __syncthreads();
    // GT(i, J) = GX2(i, J-1) + GX1(i, J): the x-sweep part of gm and gp at y-face (i, J)
    double GT[SLOTS][3];
#pragma unroll
    for (int s = 0; s < SLOTS; s++) {
        const int i = pi[s], j = pj[s];
        const int p = tid + s * NT;
        if (in[s] && i >= 1 && i <= TX && j >= 1 && j <= TY + 1) {
#pragma unroll
            for (int k = 0; k < 3; k++) GT[s][k] = E[k * NPOS + p - PX] + GA[s][k];
        } else {
#pragma unroll
            for (int k = 0; k < 3; k++) GT[s][k] = 0.0;
        }
    }
    __syncthreads();

    // ================= y-faces: normal Riemann problems =================
#pragma unroll
    for (int s = 0; s < SLOTS; s++) {
        const int i = pi[s], j = pj[s];
        double sw[3] = {0.0, 0.0, 0.0};
        double w[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
        if (in[s] && i <= TX + 1) {
            const int kL = G::ti(i, j - 1), kR = G::ti(i, j);
            // direction-local order for y: (h, hv, hu)
            rpn2_face<OWN>(T[kL], T[2 * NP2 + kL], T[NP2 + kL], T[3 * NP2 + kL],
                           T[kR], T[2 * NP2 + kR], T[NP2 + kR], T[3 * NP2 + kR], P, sw, w,
                           RightCell{CV[s], CU[s], CS[s]});
#pragma unroll
            for (int mw = 0; mw < 3; mw++) {
                const double v = dtdy * sw[mw];
                cfl = fmaxf64(cfl, v);
                cfl = fmaxf64(cfl, -v);
            }
        }
        double wc[7];
        compress(w, wc);
        Sr[s][0] = sw[0]; Sr[s][1] = sw[1]; Sr[s][2] = sw[2];
        if (in[s]) {
            const int p = tid + s * NT;
#pragma unroll
            for (int k = 0; k < 7; k++) E[k * NPOS + p] = wc[k];
        }
    }
    __syncthreads();

    // ================= y-faces: limiter + corrections =================
    double GMr[SLOTS][3], GPr[SLOTS][3];
#pragma unroll
    for (int s = 0; s < SLOTS; s++) {
        const int i = pi[s], j = pj[s];
        const int p = tid + s * NT;
#pragma unroll
        for (int k = 0; k < 3; k++) { GMr[s][k] = 0.0; GPr[s][k] = 0.0; AM[s][k] = 0.0; AP[s][k] = 0.0; }
        if (in[s] && i <= TX + 1 && j >= 1 && j <= TY + 1) {
            double wl[7], wc[7], wr[7], fm[3], fp[3];
#pragma unroll
            for (int k = 0; k < 7; k++) {
                wl[k] = E[k * NPOS + p - PX];
                wc[k] = E[k * NPOS + p];
                wr[k] = E[k * NPOS + p + PX];
            }
            face_second_order<1, DEF>(wc, Sr[s], wl, wr, dtdy, P, fm, fp, AM[s], AP[s]);
            // gm = GT + faddm in physical order (h, hu, hv) = local (0, 2, 1)
            GMr[s][0] = GT[s][0] + fm[0];
            GMr[s][1] = GT[s][1] + fm[2];
            GMr[s][2] = GT[s][2] + fm[1];
            GPr[s][0] = GT[s][0] + fp[0];
            GPr[s][1] = GT[s][1] + fp[2];
            GPr[s][2] = GT[s][2] + fp[1];
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < SLOTS; s++)
        if (in[s]) {
            const int p = tid + s * NT;
#pragma unroll
            for (int k = 0; k < 3; k++) E[k * NPOS + p] = AM[s][k];
        }
    __syncthreads();

    // ================= y-transverse (cell-centred rpt2) =================
    const double hdy = 0.5 * dtdy;
#pragma unroll
    for (int s = 0; s < SLOTS; s++) {
        const int i = pi[s], j = pj[s];
        const int p = tid + s * NT;
#pragma unroll
        for (int k = 0; k < 3; k++) { GA[s][k] = 0.0; GB[s][k] = 0.0; }
        if (transverse && in[s] && i <= TX + 1 && j >= 1 && j <= TY) {
            const int kc = G::ti(i, j);
            const Rpt2Cell C =
                OWN ? rpt2_setup_c(T[kc], CV[s], CU[s], CS[s], T[3 * NP2 + kc], T[3 * NP2 + G::ti(i - 1, j)],
                                   T[3 * NP2 + G::ti(i + 1, j)], drytol)
                    : rpt2_setup(T[kc], T[2 * NP2 + kc], T[NP2 + kc], T[3 * NP2 + kc], T[3 * NP2 + G::ti(i - 1, j)],
                                 T[3 * NP2 + G::ti(i + 1, j)], g, drytol);
            const double amn[3] = {E[p + PX], E[NPOS + p + PX], E[2 * NPOS + p + PX]};
            cell_transverse(C, amn, AP[s], hdy, GA[s], GB[s]);  // GA = left (side 1), GB = right (side 2)
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < SLOTS; s++)
        if (in[s]) {
            const int p = tid + s * NT;
#pragma unroll
            for (int k = 0; k < 3; k++) E[k * NPOS + p] = GB[s][k];
        }
    __syncthreads();
    // fm, fp at x-face (I, j) = (faddm_x + GY2(I-1, j)) + GY1(I, j), physical order
    double FMr[SLOTS][3], FPr[SLOTS][3];
#pragma unroll
    for (int s = 0; s < SLOTS; s++) {
        const int i = pi[s], j = pj[s];
        const int p = tid + s * NT;
#pragma unroll
        for (int k = 0; k < 3; k++) { FMr[s][k] = 0.0; FPr[s][k] = 0.0; }
        if (in[s] && i >= 1 && i <= TX + 1 && j >= 1 && j <= TY) {
            const double g2[3] = {E[p - 1], E[2 * NPOS + p - 1], E[NPOS + p - 1]};  // local -> physical
            const double g1[3] = {GA[s][0], GA[s][2], GA[s][1]};
#pragma unroll
            for (int k = 0; k < 3; k++) {
                if constexpr (FREG) {
                    FMr[s][k] = FX[s][k] + g2[k] + g1[k];
                    FPr[s][k] = FX[s][3 + k] + g2[k] + g1[k];
                } else {
                    const int f = G::fi_(i, j);
                    FMr[s][k] = F[k * G::NF + f] + g2[k] + g1[k];
                    FPr[s][k] = F[(3 + k) * G::NF + f] + g2[k] + g1[k];
                }
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < SLOTS; s++)
        if (in[s]) {
            const int p = tid + s * NT;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                E[k * NPOS + p] = FMr[s][k];
                E[(3 + k) * NPOS + p] = GMr[s][k];
            }
        }
    __syncthreads();

    // ================= conservative update + src2 =================
    double hm = 0.0;
    bool bad = false;
#pragma unroll
    for (int s = 0; s < SLOTS; s++) {
        const int i = pi[s], j = pj[s];
        const int p = tid + s * NT;
        if (valid && in[s] && i >= 1 && i <= TX && j >= 1 && j <= TY) {
            const int k = G::ti(i, j);
            const int pk = GP::ti(ox + i, oy + j);
            double qn[3];
#pragma unroll
            for (int m = 0; m < 3; m++) {
                const double fmr = E[m * NPOS + p + 1];
                const double gmt = E[(3 + m) * NPOS + p + PX];
                qn[m] = T[m * NP2 + k] - dtdx * (fmr - FPr[s][m]) - dtdy * (gmt - GPr[s][m]);
            }
            if (P.src_term > 0) src2_cell(qn[0], qn[1], qn[2], T[3 * NP2 + k], dt, P);
            dst[pk] = qn[0];
            dst[PNP2 + pk] = qn[1];
            dst[2 * PNP2 + pk] = qn[2];
            hm = fmaxf64(hm, qn[0]);
            bad |= !finite3(qn[0], qn[1], qn[2]);
        }
    }

    // ================= CFL and depth reductions =================
    // (the per-patch MAXCFL_CAP is applied by the caller, once per patch; a
    // non-finite new state makes the CFL number NaN, which the cap keeps)
    double v = bad ? quiet_nan() : cfl, hmx = hm;
    block_max2<NT>(v, hmx, red, tid);
    if (tid == 0) {
        cfl_res = v;
        hmax_res = hmx;
    }
}

// Tile decomposition of an MX x MY patch into TX x TY tiles.
template <int MX, int MY, int TX, int TY>
struct Tiling {
    static_assert(MX % TX == 0 && MY % TY == 0, "tiles must divide the patch");
    static constexpr int NTX = MX / TX;
    static constexpr int NTILES = NTX * (MY / TY);
    __device__ static void origin(int tile, int& ox, int& oy)
    {
        ox = (tile % NTX) * TX;
        oy = (tile / NTX) * TY;
    }
};

// Per-patch result of a tiled step: every tile folds its CFL maximum and
// depth maximum into the patch's slots (atomicMax on the bit patterns of
// non-negative values; zeros are left out because -0.0 sorts above every
// positive pattern), then counts itself in; the tile that completes the
// count takes the patch's totals and clears the slots for the next use.
// Returns true in thread 0 of that last tile, with the totals in cfl / hm.
__device__ __forceinline__ bool tile_fold(unsigned long long* tcfl, unsigned long long* th, int* tcnt,
                                          int patch, int ntiles, double& cfl, double& hm)
{
    if (cfl != 0.0) atomicMax(tcfl + patch, (unsigned long long)__double_as_longlong(cfl));
    if (hm != 0.0) atomicMax(th + patch, (unsigned long long)__double_as_longlong(hm));
    __threadfence();
    const int done = atomicAdd(tcnt + patch, 1);
    if (done != ntiles - 1) return false;
    __threadfence();
    cfl = __longlong_as_double((long long)atomicExch(tcfl + patch, 0ull));
    hm = __longlong_as_double((long long)atomicExch(th + patch, 0ull));
    atomicExch(tcnt + patch, 0);
    return true;
}

// Plain step kernel: one workgroup per tile, stored ghost cells.
template <int MX, int MY, int TX, int TY, int NT, bool FREG, bool DEF>
__global__ void __launch_bounds__(NT) k_step(StepArgs a)
{
    using G = Geo<TX, TY>;
    using TL = Tiling<MX, MY, TX, TY>;
    __shared__ double T[4 * G::NP2];
    __shared__ double E[7 * G::NPOS];
    __shared__ double F[FREG ? 1 : 6 * G::NF];
    __shared__ double red[2 * (NT / 64)];
    const int bid = blockIdx.x;
    if (bid >= a.count * TL::NTILES) return;
    const int patch = a.first + bid / TL::NTILES;
    int ox, oy;
    TL::origin(bid % TL::NTILES, ox, oy);
    const double* src = a.cur ? a.q1 : a.q0;
    double* dst = a.q1 ? (a.cur ? a.q0 : a.q1) : a.q0;
    double cfl = 0.0, hm = 0.0;
    step_patch<MX, MY, TX, TY, NT, false, DEF, FREG>(T, E, F, red, a.desc + patch, patch, ox, oy, src, dst, a.aux,
                                                     a.dt, a.P, cfl, hm, threadIdx.x, true);
    if (threadIdx.x == 0) {
        bool last = true;
        if (TL::NTILES > 1) last = tile_fold(a.tile_cfl, a.tile_h, a.tile_cnt, patch, TL::NTILES, cfl, hm);
        if (last) {
            if (cfl > 1000.0) cfl = 0.0;  // MAXCFL_CAP per patch (fc2d_cpucuda.cpp:392)
            a.cfl_out[patch] = cfl;
            if (a.hmax_out) a.hmax_out[patch] = hm;
        }
    }
}

// What an active item needs before its tile can be loaded: the patch, its
// descriptor and its scenario's buffer, dt and Manning n.  Reading them is a
// chain of dependent memory accesses (list -> descriptor -> control record),
// so the item before fetches them for its successor (ItemPrefetch) and the
// item starts with them in LDS.
struct ItemHead {
    PatchDesc d;
    int patch, cur;
    double dt, manning;
};

__device__ __forceinline__ void load_head(const BatchArgs& a, int patch, ItemHead* h)
{
    // the descriptor as 8-byte words, in registers (a struct copy would go
    // through the private stack)
    static_assert(sizeof(PatchDesc) % 8 == 0 && offsetof(PatchDesc, scen) % 8 == 0, "PatchDesc layout");
    constexpr int W = sizeof(PatchDesc) / 8;
    const long long* s = reinterpret_cast<const long long*>(a.desc + patch);
    long long w[W];
#pragma unroll
    for (int k = 0; k < W; k++) w[k] = s[k];
    long long* d = reinterpret_cast<long long*>(&h->d);
#pragma unroll
    for (int k = 0; k < W; k++) d[k] = w[k];
    const int scen = (int)(unsigned)(w[offsetof(PatchDesc, scen) / 8] & 0xffffffffll);
    const ScenCtrl& c = a.ctrl[scen];
    h->patch = patch;
    h->cur = c.cur;
    h->dt = c.dt;
    h->manning = c.manning_n;
}

// Run by one thread of the current active item: takes the next item from
// the cursor, publishes it, and when it is active fills its ItemHeads (one
// per patch of the item: entries next*NPP .. of the active tile list).
template <int NTILES, int NPP>
struct ItemPrefetch {
    const BatchArgs* pa;
    int* next_slot;
    ItemHead* next_head;
    int nwet;
    __device__ void operator()() const
    {
        const int next = atomicAdd(&pa->counters[NEXT_WET], 1);
        *next_slot = next;
#pragma unroll
        for (int g = 0; g < NPP; g++) {
            const int e = next * NPP + g;
            if (e < nwet) load_head(*pa, pa->list_wet[e / NTILES], next_head + g);
        }
    }
};

// Persistent step kernel over the device work list of active patches:
// each workgroup takes the next (patch, tile) entry from an atomic counter,
// pulls the tile's halo from the patch and its neighbours, steps it, and
// folds the patch's capped CFL number into the scenario's maximum
// (atomicMax on the bit pattern of a CFL > 0).  One entry of the persistent
// kernel.  (The library is built with machine-level loop-invariant code
// motion off: hoisting the body's constant and index materialisations out
// of the persistent loop would keep them live across the whole step and
// spill.)
//
// NPP patches per workgroup (untiled variants): the item is entries
// idx*NPP .. idx*NPP+NPP-1 of the active list, one per group of NT threads.
template <int MX, int MY, int TX, int TY, int NT, int NPP, bool FREG, bool DEF>
__device__ __forceinline__ void wet_item(const BatchArgs* pa, int idx, const ItemHead* head, ItemHead* next_head,
                                         int* next_slot, int nwet)
{
    using G = Geo<TX, TY>;
    using TL = Tiling<MX, MY, TX, TY>;
    static_assert(NPP == 1 || TL::NTILES == 1, "several patches per workgroup only for variants with one tile per patch");
    __shared__ double T[NPP * 4 * G::NP2];
    __shared__ double E[NPP * 7 * G::NPOS];
    __shared__ double F[FREG ? 1 : NPP * 6 * G::NF];
    __shared__ double red[NPP * 2 * (NT / 64)];
    const BatchArgs& a = *pa;
    const int grp = NPP > 1 ? (int)threadIdx.x / NT : 0;
    const int ltid = NPP > 1 ? (int)threadIdx.x - grp * NT : (int)threadIdx.x;
    const int entry = idx * NPP + grp;
    const bool valid = NPP == 1 || entry < nwet;
    // a group without a patch computes on a zero tile with group 0's header
    const ItemHead* hd = valid ? head + grp : head;
    const int patch = hd->patch;
    int ox, oy;
    TL::origin(entry % TL::NTILES, ox, oy);
    const int scen = hd->d.scen;
    const int cur = hd->cur;
    const double dt = hd->dt;
    const double* src = cur ? a.q1 : a.q0;
    double* dst = cur ? a.q0 : a.q1;
    Params P = a.P;
    P.manning_n = hd->manning;
    if (ltid == 0 && valid) a.dry_age[patch] = 0;
    double cfl = 0.0, hm = 0.0;
    const ItemPrefetch<TL::NTILES, NPP> pf{pa, next_slot, next_head, nwet};
    step_patch<MX, MY, TX, TY, NT, true, DEF, FREG>(T + grp * 4 * G::NP2, E + grp * 7 * G::NPOS,
                                                    F + (FREG ? 0 : grp * 6 * G::NF), red + grp * 2 * (NT / 64),
                                                    &hd->d, patch, ox, oy, src, dst, a.aux, dt, P, cfl, hm, ltid,
                                                    valid, pf);
    if (ltid == 0 && valid) {
        bool last = true;
        if (TL::NTILES > 1) last = tile_fold(a.tile_cfl, a.tile_h, a.tile_cnt, patch, TL::NTILES, cfl, hm);
        if (last) {
            if (cfl > 1000.0) cfl = 0.0;  // MAXCFL_CAP per patch (fc2d_cpucuda.cpp:392)
            a.hmax[(size_t)(cur ^ 1) * a.capacity + patch] = hm;
            // The unsigned order of bit patterns is the order of non-negative
            // doubles, but -0.0 (which the CFL maximum can return for a patch
            // with no waves) sorts above them all: zeros are left out (the
            // scenario's maximum starts at +0.0).  NaN still propagates.
            if (cfl != 0.0) atomicMax(a.cflbits + scen, (unsigned long long)__double_as_longlong(cfl));
        }
    }
    __syncthreads();
}

// Dry patches per work item of the unified kernel: about eight cells per
// thread (a whole patch at least), so that the item's loads are all in
// flight at once without crowding the registers of the active items.
template <int MX, int MY, int NT>
constexpr int dry_per_item()
{
    return 8 * NT / (MX * MY) > 0 ? 8 * NT / (MX * MY) : 1;
}

// One dry work item of the unified kernel: dry_per_item consecutive entries of
// the dry list.  A dry patch's whole tile is at most dry_tolerance, so every
// wave and transverse split vanishes and the step is b4step2 + src2 on the
// interior: a cell with h < dry_tolerance becomes (max(h,0), 0, 0) whatever
// its momentum (src2 then keeps zero momentum at zero), so only h is read for
// it; hu, hv and b are read only for h == dry_tolerance.
template <int MX, int MY, int NT>
__device__ __forceinline__ void dry_item(const BatchArgs* pa, int k, int next, int* next_slot, int* unsteady)
{
    using G = Geo<MX, MY>;
    constexpr int NP2 = G::NP2, CELLS = MX * MY;
    constexpr int DPI = dry_per_item<MX, MY, NT>();
    constexpr int PER = (DPI * CELLS + NT - 1) / NT;
    static_assert(DPI <= NT, "one thread per dry patch of an item updates its steady state");
    constexpr int PMASK = (1 << DRY_CUR_SHIFT) - 1;
    const BatchArgs& a = *pa;
    const int first = k * DPI;
    const int count = a.counters[CNT_DRY];
    // opaque thread index: keeps the item's index computations inside the
    // persistent loop instead of live (and spilled) across the active items
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const double tol = a.P.dry_tolerance;
    // entries (patch | cur << DRY_CUR_SHIFT), then h, then the stores: two
    // dependent memory round trips per item
    int ent[PER], kk[PER];
#pragma unroll
    for (int r = 0; r < PER; r++) {
        const int e = tid + NT * r;
        const int j = first + e / CELLS;
        ent[r] = (e < DPI * CELLS && j < count) ? a.list_dry[j] : -1;
        const int c = e % CELLS;
        kk[r] = G::ti(c % MX + 1, c / MX + 1);
    }
    double h0[PER];
#pragma unroll
    for (int r = 0; r < PER; r++)
        if (ent[r] >= 0) h0[r] = ((ent[r] >> DRY_CUR_SHIFT) ? a.q1 : a.q0)[(size_t)(ent[r] & PMASK) * 3 * NP2 + kk[r]];
#pragma unroll
    for (int r = 0; r < PER; r++) {
        if (ent[r] < 0) continue;
        const int pt = ent[r] & PMASK, cur = ent[r] >> DRY_CUR_SHIFT;
        const double* src = (cur ? a.q1 : a.q0) + (size_t)pt * 3 * NP2;
        double* dst = (cur ? a.q0 : a.q1) + (size_t)pt * 3 * NP2;
        double h = h0[r], hu = 0.0, hv = 0.0;
        if (h < tol) {
            h = fmaxf64(h, 0.0);
        } else {
            atomicOr(&unsteady[(tid + NT * r) / CELLS], 1);
            const ScenCtrl* c = a.ctrl + a.desc[pt].scen;
            Params P = a.P;
            P.manning_n = c->manning_n;
            hu = src[NP2 + kk[r]];
            hv = src[2 * NP2 + kk[r]];
            if (P.src_term > 0) src2_cell(h, hu, hv, a.aux[(size_t)pt * NP2 + kk[r]], c->dt, P);
            // a NaN depth lands here too: it sets the scenario's CFL maximum
            // to NaN, which stops the scenario at the controller
            if (!finite3(h, hu, hv))
                atomicMax(a.cflbits + a.desc[pt].scen, (unsigned long long)__double_as_longlong(quiet_nan()));
        }
        dst[kk[r]] = h;
        dst[NP2 + kk[r]] = hu;
        dst[2 * NP2 + kk[r]] = hv;
        if ((tid + NT * r) % CELLS == 0) a.hmax[(size_t)(cur ^ 1) * a.capacity + pt] = 0.0;
    }
    __syncthreads();
    if (threadIdx.x < DPI && first + (int)threadIdx.x < count) {
        const int e = a.list_dry[first + threadIdx.x];
        const int pt = e & PMASK, dstbuf = (e >> DRY_CUR_SHIFT) ^ 1;
        const int old = a.dry_age[pt];
        int age = old & 3;
        if (unsteady[threadIdx.x]) age = 0;
        else if (((old >> DRY_LAST_BIT) & 1) != dstbuf || age == 0) age = age < DRY_STEADY ? age + 1 : age;
        a.dry_age[pt] = age | (dstbuf << DRY_LAST_BIT);
    }
    if (threadIdx.x < DPI) unsteady[threadIdx.x] = 0;
    if (threadIdx.x == 0) *next_slot = next;
    __syncthreads();
}

// Unified persistent step of an attempt: the active patches' tiles first
// (wet_item), then the dry patches in items of dry_per_item (dry_item), from
// one atomic cursor.  The streaming dry items fill the tail that the
// compute-bound active items leave when the list runs out.
// 8x8 patches: the register budget for three waves per SIMD (the kernels
// need 170 VGPRs, two waves, without it; no spills at 167) is 8% faster at
// ten thousand active patches; larger patches keep their own allocation.
template <int MX, int MY, int TX, int TY, int NT, int NPP, bool FREG, bool DEF>
__global__ void __launch_bounds__(NT * NPP, (MX == 8 ? 3 : 1)) k_all(const BatchArgs* pa)
{
    using TL = Tiling<MX, MY, TX, TY>;
    constexpr int NB = NT * NPP;  // threads of the workgroup
    constexpr int DPI = dry_per_item<MX, MY, NB>();
    __shared__ int item[2];
    __shared__ ItemHead head[2][NPP];
    __shared__ int unsteady[DPI];  // per dry patch of an item: a cell at dry_tolerance
    const int nwet = pa->counters[CNT_WET] * TL::NTILES;  // active tiles
    const int witems = (nwet + NPP - 1) / NPP;             // active items
    if (threadIdx.x == 0) {
        item[0] = atomicAdd(&pa->counters[NEXT_WET], 1);
        for (int g = 0; g < NPP; g++)
            if (item[0] * NPP + g < nwet) load_head(*pa, pa->list_wet[(item[0] * NPP + g) / TL::NTILES], &head[0][g]);
    }
    if (threadIdx.x < DPI) unsteady[threadIdx.x] = 0;
    __syncthreads();
    const int limit = witems + (pa->counters[CNT_DRY] + DPI - 1) / DPI;
    for (int it = 0;; it++) {
        const int idx = item[it & 1];
        if (idx >= limit) break;
        if (idx < witems) {
            // active items come first in the cursor's order, so an active
            // item's successor's heads are always prefetched by an active item
            wet_item<MX, MY, TX, TY, NT, NPP, FREG, DEF>(pa, idx, head[it & 1], head[(it + 1) & 1],
                                                         &item[(it + 1) & 1], nwet);
        } else {
            int next = 0;
            if (threadIdx.x == 0) next = atomicAdd(&pa->counters[NEXT_WET], 1);
            dry_item<MX, MY, NB>(pa, idx - witems, next, &item[(it + 1) & 1], unsteady);
        }
    }
}

// ---------------------------------------------------------------------------
// Ghost fill on the device for same-level patches: each ghost cell takes the
// value of the cell that ForestClaw's face copy, corner copy and bc2 pass
// (left, right, bottom, top over full extents) would leave there.
// ---------------------------------------------------------------------------

template <int MX, int MY>
__global__ void __launch_bounds__(256) k_ghost(GhostArgs a)
{
    using G = Geo<MX, MY>;
    constexpr int NXP = G::NXP, NP2 = G::NP2;
    const int bid = blockIdx.x;
    if (bid >= a.count) return;
    const int patch = a.first + bid;
    const PatchDesc* dp = a.desc + patch;
    int cur = a.cur;
    if (a.ctrl) {
        const int scen = dp->scen;
        if (scen < 0) return;
        const ScenCtrl* c = a.ctrl + scen;
        if (!c->active) return;
        cur = c->cur;
    }
    double* base = cur ? a.q1 : a.q0;
    double* q = base + (size_t)patch * 3 * NP2;

    for (int c = threadIdx.x; c < NP2; c += blockDim.x) {
        const int fi = c % NXP - 1, fj = c / NXP - 1;  // Fortran indices
        if (fi >= 1 && fi <= MX && fj >= 1 && fj <= MY) continue;
        int sp, si, sj;
        bool neg_u, neg_v;
        if (!ghost_source<MX, MY>(dp->nbr, patch, fi, fj, sp, si, sj, neg_u, neg_v)) continue;
        const double* s = base + (size_t)sp * 3 * NP2;
        const int k = G::ti(si, sj);
        const double h = s[k], hu = s[NP2 + k], hv = s[2 * NP2 + k];
        q[c] = h;
        q[NP2 + c] = neg_u ? -hu : hu;
        q[2 * NP2 + c] = neg_v ? -hv : hv;
    }
}

// ---------------------------------------------------------------------------
// Hydrograph inflow patches.  The ghost ring of each listed patch is rebuilt
// in LDS in ForestClaw's order -- same-level face/corner copies (and the
// closed-form wall/extrapolation states) first, then the bc2 passes left,
// right, bottom, top over full extents (geoflood_bc2_fort.f90:7-499) -- and
// written back to the patch's own ghost cells and aux, where the step reads
// the inflow sides as stored ghosts.  The passes run in sequence because a
// later side's corner cells read what an earlier side wrote.
// ---------------------------------------------------------------------------

// One bc2 side on the LDS tile L (h, hu, hv, b planes): wall (mthbc 3),
// extrapolation (mthbc 1) or hydrograph inflow (mthbc 0).  Reads interior
// rows/columns and, for the y sides, the ghost columns the x sides wrote;
// writes only this side's ghost cells.
template <int MX, int MY>
__device__ void bc2_side(double* L, int side, int code, const PatchDesc* __restrict__ dp,
                         const ScenCtrl* __restrict__ c, double hu0, const Params& P)
{
    using G = Geo<MX, MY>;
    constexpr int NP2 = G::NP2;
    const bool xs = side < 2;
    const int n = xs ? MY : MX;
    const int items = (n + 2 * MBC) * MBC;
    for (int it = threadIdx.x; it < items; it += blockDim.x) {
        const int r = it / MBC + 1 - MBC;  // row (x sides) or column (y sides)
        const int kb = it % MBC + 1;       // ghost layer
        int gi, gj, mi, mj, bi, bj;        // ghost, mirror source, boundary cell
        if (side == 0) { gi = 1 - kb; gj = r; mi = kb; mj = r; bi = 1; bj = r; }
        else if (side == 1) { gi = MX + kb; gj = r; mi = MX + 1 - kb; mj = r; bi = MX; bj = r; }
        else if (side == 2) { gi = r; gj = 1 - kb; mi = r; mj = kb; bi = r; bj = 1; }
        else { gi = r; gj = MY + kb; mi = r; mj = MY + 1 - kb; bi = r; bj = MY; }
        const int gk = G::ti(gi, gj), mk = G::ti(mi, mj), bk = G::ti(bi, bj);
        const int nrm = xs ? 1 : 2;  // plane of the normal momentum
        if (code == NB_EXTRAP) {
            L[gk] = L[bk];
            L[NP2 + gk] = L[NP2 + bk];
            L[2 * NP2 + gk] = L[2 * NP2 + bk];
            L[3 * NP2 + gk] = L[3 * NP2 + bk];
            continue;
        }
        bool mirror = true;
        if (code == NB_INFLOW) {
            const double lo = xs ? dp->ylo : dp->xlo;
            const double del = xs ? dp->dy : dp->dx;
            const double c0 = xs ? c->hyd_y0 : c->hyd_x0;
            const double coord = lo + ((double)r - 0.5) * del;
            if (fabs(coord - c0) <= c->hyd_b / 2.0) {
                const double h1 = L[bk], hu1 = L[NP2 + bk];
                mirror = false;
                double h0;
                if (h1 < P.dry_tolerance) {
                    h0 = fmaxf64(pow_2_3(hu0 / sqrt(P.grav)), 0.001);
                } else if (hu0 != 0.0) {
                    h0 = riemann_invariant_depth(hu0, h1, hu1, P.grav);
                    if (h0 > h1) h0 = two_shock_depth(hu0, h1, hu1, P.grav);
                } else {
                    mirror = true;
                    h0 = 0.0;
                }
                if (!mirror) {
                    L[gk] = h0;
                    L[NP2 + gk] = hu0;  // the discharge goes to q(2) on every side
                    L[2 * NP2 + gk] = 0.0;
                }
            }
        }
        if (mirror) {
            L[gk] = L[mk];
            L[NP2 + gk] = L[NP2 + mk];
            L[2 * NP2 + gk] = L[2 * NP2 + mk];
            L[nrm * NP2 + gk] = -L[nrm * NP2 + mk];
            L[3 * NP2 + gk] = L[3 * NP2 + mk];
        }
    }
}

template <int MX, int MY>
__global__ void __launch_bounds__(256) k_inflow(InflowArgs a)
{
    using G = Geo<MX, MY>;
    constexpr int NXP = G::NXP, NP2 = G::NP2;
    __shared__ double L[4 * NP2];
    const int lo = a.range ? a.range[0] : 0, n = a.range ? a.range[1] : *a.list_n;
    for (int k = lo + (int)blockIdx.x; k < n; k += gridDim.x) {
        const int patch = a.list[k];
        if (patch < a.first || patch >= a.last) continue;
        const PatchDesc* dp = a.desc + patch;
        const int scen = dp->scen;
        if (scen < 0) continue;
        const ScenCtrl* c = a.ctrl + scen;
        int cur;
        double t;
        if (a.buf < 0) {
            if (!c->active) continue;
            cur = c->cur;
            t = c->t;
        } else {
            cur = a.buf;
            t = a.t;
        }
        double* base = cur ? a.q1 : a.q0;
        double* q = base + (size_t)patch * 3 * NP2;
        double* ax = a.aux + (size_t)patch * NP2;
        __syncthreads();  // the previous patch's write-back has read L
        // bc2's side passes read the ghost ring and the MBC interior rows and
        // columns next to it (mirror and boundary cells): the deeper interior
        // is never read, and not loaded
        for (int e = threadIdx.x; e < NP2; e += blockDim.x) {
            const int fi = e % NXP - 1, fj = e / NXP - 1;
            if (fi > MBC && fi <= MX - MBC && fj > MBC && fj <= MY - MBC) continue;
            const double* src = q;
            int kk = e;
            bool nu = false, nv = false;
            if (!(fi >= 1 && fi <= MX && fj >= 1 && fj <= MY)) {
                int sp, si, sj;
                if (ghost_source<MX, MY>(dp->nbr, patch, fi, fj, sp, si, sj, nu, nv)) {
                    src = base + (size_t)sp * 3 * NP2;
                    kk = G::ti(si, sj);
                }
            }
            const double h = src[kk], hu = src[NP2 + kk], hv = src[2 * NP2 + kk];
            L[e] = h;
            L[NP2 + e] = nu ? -hu : hu;
            L[2 * NP2 + e] = nv ? -hv : hv;
            L[3 * NP2 + e] = ax[e];
        }
        __syncthreads();
        // the discharge per unit width at time t, once for the patch, when a
        // ghost cell of its hydrograph side lies in the channel (bc2_side's
        // own test; every wave takes the same branch and value)
        int in_channel = 0;
        for (int side = 0; side < 4; side++) {
            if (dp->nbr[side] != NB_INFLOW) continue;
            const bool xs = side < 2;
            const int r = (int)threadIdx.x + 1 - MBC;  // one row (x sides) or column (y sides) per thread
            if (r <= (xs ? MY : MX) + MBC) {
                const double lo = xs ? dp->ylo : dp->xlo;
                const double del = xs ? dp->dy : dp->dx;
                const double c0 = xs ? c->hyd_y0 : c->hyd_x0;
                in_channel |= fabs(lo + ((double)r - 0.5) * del - c0) <= c->hyd_b / 2.0;
            }
        }
        const double hu0 = __syncthreads_or(in_channel)
                               ? hydro_interpolate(c->hyd, c->hyd + c->hyd_rows, c->hyd_rows, t)
                               : 0.0;
        for (int side = 0; side < 4; side++) {
            const int code = dp->nbr[side];
            if (code == NB_WALL || code == NB_EXTRAP || code == NB_INFLOW) bc2_side<MX, MY>(L, side, code, dp, c, hu0, a.P);
            __syncthreads();
        }
        for (int e = threadIdx.x; e < NP2; e += blockDim.x) {
            const int fi = e % NXP - 1, fj = e / NXP - 1;
            if (fi >= 1 && fi <= MX && fj >= 1 && fj <= MY) continue;
            q[e] = L[e];
            q[NP2 + e] = L[NP2 + e];
            q[2 * NP2 + e] = L[2 * NP2 + e];
            ax[e] = L[3 * NP2 + e];
        }
    }
}

// ---------------------------------------------------------------------------
// Device-side dt controller (fc2d_geoclaw_run.c:504-565): one workgroup per
// scenario reduces the patches' capped CFL numbers, then either accepts the
// step (flip the q buffer, t += dt) or rejects it (retake), and sets the next
// dt = dt * desired_cfl / maxcfl.
// ---------------------------------------------------------------------------
// One decision of the outstyle_3 loop (fc2d_geoclaw_run.c:529-580) for a
// scenario whose attempt gave CFL number m: retake with a smaller dt when m
// exceeds max_cfl, else accept (t += dt, flip the buffer) and choose the next
// dt.  A non-finite m (a NaN state or wave speed) stops the scenario with
// status SCEN_NONFINITE and leaves t, dt and the accepted buffer as they
// were; the reference has no such check.
__device__ __forceinline__ void control_step(ScenCtrl* c, double m)
{
    c->maxcfl = m;
    c->attempts += 1;
    if (!cfl_finite(m)) {
        c->status = SCEN_NONFINITE;
        c->active = 0;
    } else if (m > c->max_cfl) {
        c->retakes += 1;
        c->dt = c->dt * c->desired_cfl / m;
    } else {
        c->t = c->t + c->dt;
        c->last_dt = c->dt;
        c->cur ^= 1;
        c->nsteps += 1;
        c->dt = next_dt(c->dt, c->desired_cfl, m, c->dt_max);
        if ((c->target_steps > 0 && c->nsteps >= c->target_steps) || (c->t_end > 0.0 && c->t >= c->t_end))
            c->active = 0;
    }
}

__global__ void __launch_bounds__(256) k_control(ScenCtrl* ctrl, int nscen, const double* cfl)
{
    const int s = blockIdx.x;
    if (s >= nscen) return;
    ScenCtrl* c = ctrl + s;
    if (!c->active) return;
    __shared__ double red[4];
    double v = 0.0;
    for (int k = threadIdx.x; k < c->npatch; k += blockDim.x) v = max_nan(v, cfl[c->first_patch + k]);
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int k = 1; k < (int)(blockDim.x >> 6); k++) m = max_nan(m, red[k]);
        control_step(c, m);
    }
}

// ---------------------------------------------------------------------------
// Batched path: classification, dry-patch stream, controller.
// ---------------------------------------------------------------------------

// A patch is active for the next attempt when its tile (interior + 2-cell
// halo from the 8 neighbours) may hold a cell with h > dry_tolerance: its own
// or a neighbour's maximum interior depth exceeds the tolerance, or a side's
// ghost cells come from the host.  Active patches go to the wet list (full
// step), the others to the dry list.  Appends are aggregated per wavefront.
__global__ void __launch_bounds__(256) k_classify(const BatchArgs* pa)
{
    const BatchArgs& a = *pa;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    int kind = -1;  // -1: not stepping, 0: dry, 1: active
    int cur = 0;
    if (p < a.n) {
        const PatchDesc d = a.desc[p];
        if (d.scen >= 0) {
            const ScenCtrl* c = a.ctrl + d.scen;
            if (c->active) {
                cur = c->cur;
                const double* hm = a.hmax + (size_t)c->cur * a.capacity;
                const double tol = a.P.dry_tolerance;
                bool w = hm[p] > tol;
#pragma unroll
                for (int s = 0; s < 8; s++) {
                    const int nb = d.nbr[s];
                    if (nb >= 0) w = w || (hm[nb] > tol);
                    else if (nb == NB_STORED || nb == NB_INFLOW) w = true;
                }
                kind = w ? 1 : 0;
            }
        }
    }
    if (kind == 0 && (a.dry_age[p] & 3) >= DRY_STEADY) {
        // steady dry: the step would rewrite the same values; only the depth
        // maximum of the next state is recorded
        a.hmax[(size_t)(cur ^ 1) * a.capacity + p] = 0.0;
        kind = 2;
    }
    const unsigned long long mw = __ballot(kind == 1);
    const unsigned long long md = __ballot(kind == 0);
    const unsigned long long ms = __ballot(kind == 2);
    const int lane = threadIdx.x & 63;
    int bw = 0, bd = 0;
    if (lane == 0) {
        if (mw) bw = atomicAdd(&a.counters[CNT_WET], __popcll(mw));
        if (md) bd = atomicAdd(&a.counters[CNT_DRY], __popcll(md));
        if (ms) atomicAdd(&a.counters[CNT_SKIP], __popcll(ms));
    }
    bw = __shfl(bw, 0, 64);
    bd = __shfl(bd, 0, 64);
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    if (kind == 1) a.list_wet[bw + __popcll(mw & below)] = p;
    if (kind == 0) a.list_dry[bd + __popcll(md & below)] = p | (cur << DRY_CUR_SHIFT);
}

// Controller of the batched path: the per-scenario CFL maximum was folded by
// atomicMax in the step kernels; apply the dt rule (fc2d_geoclaw_run.c:
// 529-580), clear the maxima and the list counters for the next attempt.
__global__ void __launch_bounds__(64) k_batch_control(const BatchArgs* pa)
{
    const BatchArgs& a = *pa;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s == 0) {
        a.stats[0] += a.counters[CNT_WET];
        a.stats[1] += a.counters[CNT_DRY] + a.counters[CNT_SKIP];
        a.stats[2] += 1;
        a.stats[3] += a.counters[CNT_SKIP];
        for (int k = 0; k < NCOUNTERS; k++) a.counters[k] = 0;
    }
    if (s >= a.nscen) return;
    ScenCtrl* c = a.ctrl + s;
    if (!c->active) return;
    const double m = __longlong_as_double((long long)a.cflbits[s]);
    a.cflbits[s] = 0ull;
    control_step(c, m);
}

// Maximum interior depth of patches [first, first+n) of a q buffer.
template <int MX, int MY>
__global__ void __launch_bounds__(256) k_hmax(const double* q, double* hmax, int first, int n)
{
    using G = Geo<MX, MY>;
    __shared__ double red[4];
    const int b = blockIdx.x;
    if (b >= n) return;
    const double* src = q + (size_t)(first + b) * 3 * G::NP2;
    double m = 0.0;
    for (int e = threadIdx.x; e < MX * MY; e += 256) m = fmaxf64(m, src[G::ti(e % MX + 1, e / MX + 1)]);
    m = block_max<256>(m, red, threadIdx.x);
    if (threadIdx.x == 0) hmax[first + b] = m;
}

// ---------------------------------------------------------------------------
// ASCII frame output, device part (fc2d_geoclaw_fort_write_file,
// fc2d_geoclaw_output_ascii_fort.f:83-108): per interior cell in row order
// (j outer, i inner) h, hu, hv with |value| < 1e-99 set to +0 -- in the
// state as well, as the reference does to the patch it writes -- and
// eta = h + b, 0 when |eta| < 1e-99.  out holds n patches of MX*MY*4 values.
// ---------------------------------------------------------------------------
template <int MX, int MY>
__global__ void __launch_bounds__(256) k_frame_pack(double* q, const double* aux, int first, int n, double* out)
{
    using G = Geo<MX, MY>;
    const int b = blockIdx.x;
    if (b >= n) return;
    double* qp = q + (size_t)(first + b) * 3 * G::NP2;
    const double* ap = aux + (size_t)(first + b) * G::NP2;
    double* o = out + (size_t)b * MX * MY * 4;
    for (int c = threadIdx.x; c < MX * MY; c += 256) {
        const int k = G::ti(c % MX + 1, c / MX + 1);
        double v[3];
#pragma unroll
        for (int m = 0; m < 3; m++) {
            v[m] = qp[m * G::NP2 + k];
            if (fabs(v[m]) < 1e-99) {
                v[m] = 0.0;
                qp[m * G::NP2 + k] = 0.0;
            }
        }
        double eta = v[0] + ap[k];
        if (fabs(eta) < 1e-99) eta = 0.0;
        o[4 * c] = v[0];
        o[4 * c + 1] = v[1];
        o[4 * c + 2] = v[2];
        o[4 * c + 3] = eta;
    }
}

// Gauge sample (fc2d_geoclaw_update_gauge, geoclaw_gauges_fort.f90:1-77):
// bilinear in the four cells around the gauge unless one of their depths is
// below dry_tolerance/10, then the containing cell; |q| < 1e-99 -> 0.  Cells
// outside the interior (a gauge in the last row or column) take the value a
// ghost fill would give them: the neighbour's interior, the wall mirror or
// the extrapolation, or the stored ghost cell.  b is the patch's aux.
template <int MX, int MY>
__device__ __forceinline__ double gauge_cell(const GaugeArgs& a, int patch, int fi, int fj, int m)
{
    using G = Geo<MX, MY>;
    int sp = patch, si = fi, sj = fj;
    bool nu = false, nv = false;
    const bool interior = fi >= 1 && fi <= MX && fj >= 1 && fj <= MY;
    if (!interior && !ghost_source<MX, MY>(a.desc[patch].nbr, patch, fi, fj, sp, si, sj, nu, nv)) {
        sp = patch; si = fi; sj = fj;
    }
    const double v = a.q[(size_t)sp * 3 * G::NP2 + m * G::NP2 + G::ti(si, sj)];
    return (m == 1 && nu) || (m == 2 && nv) ? -v : v;
}

template <int MX, int MY>
__global__ void __launch_bounds__(64) k_gauges(GaugeArgs a)
{
    using G = Geo<MX, MY>;
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.n) return;
    const int patch = a.slot[g];
    const PatchDesc& d = a.desc[patch];
    const double xc = a.xy[2 * g], yc = a.xy[2 * g + 1];
    const int ii = (int)((xc - d.xlo) / d.dx) + 1;
    const int jj = (int)((yc - d.ylo) / d.dy) + 1;
    const double xcent = d.xlo + (ii - .5) * d.dx;
    const double ycent = d.ylo + (jj - .5) * d.dy;
    const double xoff = (xc - xcent) / d.dx;
    const double yoff = (yc - ycent) / d.dy;
    const double drytol2 = 0.1 * a.dry_tolerance;
    const double* ap = a.aux + (size_t)patch * G::NP2;
    double qv[3], topo;
    if (gauge_cell<MX, MY>(a, patch, ii, jj, 0) < drytol2 || gauge_cell<MX, MY>(a, patch, ii + 1, jj, 0) < drytol2 ||
        gauge_cell<MX, MY>(a, patch, ii, jj + 1, 0) < drytol2 ||
        gauge_cell<MX, MY>(a, patch, ii + 1, jj + 1, 0) < drytol2) {
#pragma unroll
        for (int m = 0; m < 3; m++) qv[m] = gauge_cell<MX, MY>(a, patch, ii, jj, m);
        topo = ap[G::ti(ii, jj)];
    } else {
#pragma unroll
        for (int m = 0; m < 3; m++)
            qv[m] = (1.0 - xoff) * (1.0 - yoff) * gauge_cell<MX, MY>(a, patch, ii, jj, m) +
                    xoff * (1.0 - yoff) * gauge_cell<MX, MY>(a, patch, ii + 1, jj, m) +
                    (1.0 - xoff) * yoff * gauge_cell<MX, MY>(a, patch, ii, jj + 1, m) +
                    xoff * yoff * gauge_cell<MX, MY>(a, patch, ii + 1, jj + 1, m);
        topo = (1.0 - xoff) * (1.0 - yoff) * ap[G::ti(ii, jj)] + xoff * (1.0 - yoff) * ap[G::ti(ii + 1, jj)] +
               (1.0 - xoff) * yoff * ap[G::ti(ii, jj + 1)] + xoff * yoff * ap[G::ti(ii + 1, jj + 1)];
    }
#pragma unroll
    for (int m = 0; m < 3; m++) {
        if (fabs(qv[m]) < 1e-99) qv[m] = 0.0;
        a.out[4 * g + m] = qv[m];
    }
    a.out[4 * g + 3] = topo;
}

// ---------------------------------------------------------------------------
// Regrid between levels (refinement ratio 2), one workgroup per (family,
// child).  Coarse cells outside the coarse patch's interior are read through
// ghost_source, as the halo pull does.
// ---------------------------------------------------------------------------

// One coarse cell from four fine cells k[0..3] (x fastest), the per-cell rule
// of fc2d_geoclaw_average_fort.f (average2coarse :482-603, average_face
// :33-320, average_corner :323-480) with mcapa = 0: the wet-cell mean of eta,
// the mean depth capped by it over the coarse bathymetry bc, momentum
// rescaled.
template <int MX, int MY>
__device__ __forceinline__ void average_four(const double* __restrict__ qf, const double* __restrict__ af,
                                             const int k[4], double bc, double drytol, double& hc, double& huc,
                                             double& hvc)
{
    using G = Geo<MX, MY>;
    double etasum = 0.0, hsum = 0.0, husum = 0.0, hvsum = 0.0;
    int nwet = 0;
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const double hf = qf[k[m]] * 1.0;
        const double bf = af[k[m]] * 1.0;
        double huf = qf[G::NP2 + k[m]] * 1.0;
        double hvf = qf[2 * G::NP2 + k[m]] * 1.0;
        double etaf;
        if (hf > drytol) {
            etaf = hf + bf;
            nwet++;
        } else {
            etaf = 0.0;
            huf = 0.0;
            hvf = 0.0;
        }
        hsum = hsum + hf;
        husum = husum + huf;
        hvsum = hvsum + hvf;
        etasum = etasum + etaf;
    }
    hc = 0.0;
    huc = 0.0;
    hvc = 0.0;
    if (nwet > 0) {
        const double etaav = etasum / (double)nwet;
        const double hav = hsum / (double)nwet;
        const double e = etaav - bc * 1.0;
        const double hmx = e > 0.0 ? e : 0.0;
        hc = hav < hmx ? hav : hmx;
        huc = (hc / hsum) * husum;
        hvc = (hc / hsum) * hvsum;
    }
}

// fc2d_geoclaw_fort_average2coarse (fc2d_geoclaw_average_fort.f:482-603),
// mcapa = 0: quadrant g of the coarse interior from child g.
template <int MX, int MY>
__global__ void __launch_bounds__(256) k_average2coarse(RegridArgs a)
{
    using G = Geo<MX, MY>;
    const int f = blockIdx.x >> 2, gch = blockIdx.x & 3;
    if (f >= a.n) return;
    const int ig = gch % 2, jg = gch / 2;
    const double* qf = a.q + (size_t)a.fs[4 * f + gch] * 3 * G::NP2;
    const double* af = a.aux + (size_t)a.fs[4 * f + gch] * G::NP2;
    double* qc = a.q + (size_t)a.cs[f] * 3 * G::NP2;
    const double* ac = a.aux + (size_t)a.cs[f] * G::NP2;
    for (int c = threadIdx.x; c < (MX / 2) * (MY / 2); c += 256) {
        const int i = c % (MX / 2) + 1, j = c / (MX / 2) + 1;
        const int i1 = i + ig * MX / 2, j1 = j + jg * MY / 2;
        int k[4];
#pragma unroll
        for (int m = 0; m < 4; m++) k[m] = G::ti((i - 1) * 2 + (m & 1) + 1, (j - 1) * 2 + (m >> 1) + 1);
        const int k1 = G::ti(i1, j1);
        double hc, huc, hvc;
        average_four<MX, MY>(qf, af, k, ac[k1], a.dry_tolerance, hc, huc, hvc);
        qc[k1] = hc / 1.0;
        qc[G::NP2 + k1] = huc / 1.0;
        qc[2 * G::NP2 + k1] = hvc / 1.0;
    }
}

// ForestClaw's fclaw2d_clawpatch_compute_slopes at mth = 5 (the slope of
// fc2d_geoclaw_fort_interpolate2fine): the smallest of 2|sl|, 2|sr|, |sc|
// with the sign of sc = (sl + sr)/2, 0 at an extremum.
__device__ __forceinline__ double amr_slope(double sl, double sr)
{
    const double sc = (sl + sr) / 2.0;
    const double a = 2 * fabs(sl), b = 2 * fabs(sr), c = fabs(sc);
    double v = a < b ? a : b;
    v = v < c ? v : c;
    const double mn = sl < sr ? sl : sr, mxv = sl > sr ? sl : sr;
    if (mn <= 0 && mxv >= 0) return 0.0;
    return copysign(v, sc);
}

template <int MX, int MY>
__device__ __forceinline__ double regrid_cell(const RegridArgs& a, int patch, int fi, int fj, int m)
{
    using G = Geo<MX, MY>;
    int sp = patch, si = fi, sj = fj;
    bool nu = false, nv = false;
    const bool interior = fi >= 1 && fi <= MX && fj >= 1 && fj <= MY;
    if (!interior && !ghost_source<MX, MY>(a.desc[patch].nbr, patch, fi, fj, sp, si, sj, nu, nv)) {
        sp = patch; si = fi; sj = fj;
    }
    const double v = a.q[(size_t)sp * 3 * G::NP2 + m * G::NP2 + G::ti(si, sj)];
    return (m == 1 && nu) || (m == 2 && nv) ? -v : v;
}

// The limited linear reconstruction of coarse cell (ic, jc) of patch cp
// (fc2d_geoclaw_interpolate_fort.f: interpolate2fine :448-505, interpolate_face
// :117-171, interpolate_corner :351-396), sea_level = 0: qc, slopes gx, gy of
// eta (dry-limited 3x3 stencil) and of hu, hv.
template <int MX, int MY>
__device__ __forceinline__ void reconstruct_coarse(const RegridArgs& a, int cp, const double* __restrict__ ac,
                                                   int ic, int jc, double qc[3], double gx[3], double gy[3])
{
    using G = Geo<MX, MY>;
    double eb[3][3];
#pragma unroll
    for (int ii = -1; ii <= 1; ii++)
#pragma unroll
        for (int jj = -1; jj <= 1; jj++) {
            const double h = regrid_cell<MX, MY>(a, cp, ic + ii, jc + jj, 0);
            const double b = ac[G::ti(ic + ii, jc + jj)];
            eb[ii + 1][jj + 1] = h < a.dry_tolerance ? 0.0 : h + b;
        }
    qc[0] = eb[1][1];
    gx[0] = amr_slope(qc[0] - eb[0][1], eb[2][1] - qc[0]);
    gy[0] = amr_slope(qc[0] - eb[1][0], eb[1][2] - qc[0]);
#pragma unroll
    for (int mq = 1; mq < 3; mq++) {
        qc[mq] = regrid_cell<MX, MY>(a, cp, ic, jc, mq);
        gx[mq] = amr_slope(qc[mq] - regrid_cell<MX, MY>(a, cp, ic - 1, jc, mq),
                           regrid_cell<MX, MY>(a, cp, ic + 1, jc, mq) - qc[mq]);
        gy[mq] = amr_slope(qc[mq] - regrid_cell<MX, MY>(a, cp, ic, jc - 1, mq),
                           regrid_cell<MX, MY>(a, cp, ic, jc + 1, mq) - qc[mq]);
    }
}

// fc2d_geoclaw_fort_interpolate2fine (fc2d_geoclaw_interpolate_fort.f:
// 405-579), sea_level = 0: child g's interior and the ghost layers on the
// family's inner sides from the coarse patch.
template <int MX, int MY>
__global__ void __launch_bounds__(256) k_interpolate2fine(RegridArgs a)
{
    using G = Geo<MX, MY>;
    const int f = blockIdx.x >> 2, gch = blockIdx.x & 3;
    if (f >= a.n) return;
    const int ig = gch % 2, jg = gch / 2;
    const int cp = a.cs[f];
    const double* ac = a.aux + (size_t)cp * G::NP2;
    double* qf = a.q + (size_t)a.fs[4 * f + gch] * 3 * G::NP2;
    const double* af = a.aux + (size_t)a.fs[4 * f + gch] * G::NP2;
    constexpr int NI = MX / 2 + 1, NJ = MY / 2 + 1;
    for (int c = threadIdx.x; c < NI * NJ; c += 256) {
        const int i = (1 - ig) + c % NI, j = (1 - jg) + c / NI;
        const int ic = i + ig * MX / 2, jc = j + jg * MY / 2;
        double qc[3], gx[3], gy[3];
        reconstruct_coarse<MX, MY>(a, cp, ac, ic, jc, qc, gx, gy);
#pragma unroll
        for (int mq = 0; mq < 3; mq++)
#pragma unroll
            for (int ii = 1; ii <= 2; ii++)
#pragma unroll
                for (int jj = 1; jj <= 2; jj++) {
                    const double shiftx = (ii - 2 / 2.0 - 0.5) / 2;
                    const double shifty = (jj - 2 / 2.0 - 0.5) / 2;
                    const int k = G::ti((i - 1) * 2 + ii, (j - 1) * 2 + jj);
                    const double v = qc[mq] + shiftx * gx[mq] + shifty * gy[mq];
                    qf[mq * G::NP2 + k] = mq == 0 ? v - af[k] : v;
                }
    }
}

// ---------------------------------------------------------------------------
// Coarse/fine ghost transfers, one wave per link l between coarse slot cs[l]
// and its half-size neighbour fs[l] of the same orientation across side[l]:
// 0..3 = face 0..3 of the coarse patch (left, right, bottom, top), the
// neighbour in half igrid[l] along it; 4..7 = corner 0..3 (lower-left,
// lower-right, upper-left, upper-right).  The neighbour covers coarse cells
// ox+1..ox+MX/2 by oy+1..oy+MY/2, so coarse cell (ic, jc) is fine cells
// 2(ic-ox)-1 + {0,1} by 2(jc-oy)-1 + {0,1}.
// ---------------------------------------------------------------------------
template <int MX, int MY>
__device__ __forceinline__ void half_origin(int side, int igrid, int& ox, int& oy)
{
    ox = (side == 0 || side == 4 || side == 6) ? -MX / 2 : (side == 1 || side == 5 || side == 7) ? MX : igrid * MX / 2;
    oy = (side == 2 || side == 4 || side == 5) ? -MY / 2 : (side == 3 || side == 6 || side == 7) ? MY : igrid * MY / 2;
}

// fc2d_geoclaw_fort_average_face / _average_corner (fc2d_geoclaw_average_fort.f:
// 33-480), mcapa = 0: the coarse ghost cells the fine neighbour covers.
// The buffer of link l: a.q, or in the resident mode the current buffer of
// the coarse patch's scenario (nullptr: an inactive scenario, skipped).
__device__ __forceinline__ double* link_buffer(const RegridArgs& a, int l)
{
    if (!a.ctrl) return a.q;
    const int sc = a.desc[a.cs[l]].scen;
    if (sc < 0) return nullptr;
    const ScenCtrl& c = a.ctrl[sc];
    if (!c.active) return nullptr;
    return c.cur ? a.q1 : a.q;
}

template <int MX, int MY>
__device__ __forceinline__ void average_link(const RegridArgs& a, int l, double* q);
template <int MX, int MY>
__device__ __forceinline__ void interpolate_link(const RegridArgs& a, int l);

template <int MX, int MY>
__global__ void __launch_bounds__(64) k_average_ghost(RegridArgs a)
{
    const int lo = a.range ? a.range[0] : 0, hi = a.range ? a.range[1] : a.n;
    for (int l = lo + (int)blockIdx.x; l < hi; l += gridDim.x) {
        double* q = link_buffer(a, l);
        if (q) average_link<MX, MY>(a, l, q);
    }
}

template <int MX, int MY>
__device__ __forceinline__ void average_link(const RegridArgs& a, int l, double* q)
{
    using G = Geo<MX, MY>;
    const int side = a.side[l], igrid = a.igrid[l];
    int ox, oy;
    half_origin<MX, MY>(side, igrid, ox, oy);
    const double* qf = q + (size_t)a.fs[l] * 3 * G::NP2;
    const double* af = a.aux + (size_t)a.fs[l] * G::NP2;
    double* qc = q + (size_t)a.cs[l] * 3 * G::NP2;
    const double* ac = a.aux + (size_t)a.cs[l] * G::NP2;
    const int ncell = side >= 4 ? MBC * MBC : MBC * (side < 2 ? MY : MX);
    for (int c = threadIdx.x; c < ncell; c += 64) {
        int ic, jc;
        if (side >= 4) {
            const int ibc = c / MBC + 1, jbc = c % MBC + 1;
            ic = (side - 4) & 1 ? MX + ibc : 1 - ibc;
            jc = (side - 4) & 2 ? MY + jbc : 1 - jbc;
        } else if (side < 2) {
            const int ibc = c % MBC + 1;
            jc = c / MBC + 1;
            ic = side == 0 ? 1 - ibc : MX + ibc;
        } else {
            const int jbc = c / MX + 1;
            ic = c % MX + 1;
            jc = side == 2 ? 1 - jbc : MY + jbc;
        }
        const int i0 = 2 * (ic - ox) - 1, j0 = 2 * (jc - oy) - 1;
        // fclaw2d_clawpatch_is_valid_average: the four cells inside the fine interior
        if (side < 4 && !(i0 >= 1 && i0 + 1 <= MX && j0 >= 1 && j0 + 1 <= MY)) continue;
        int k[4];
#pragma unroll
        for (int m = 0; m < 4; m++) k[m] = G::ti(i0 + (m & 1), j0 + (m >> 1));
        const int k1 = G::ti(ic, jc);
        double hc, huc, hvc;
        average_four<MX, MY>(qf, af, k, ac[k1], a.dry_tolerance, hc, huc, hvc);
        qc[k1] = hc / 1.0;
        qc[G::NP2 + k1] = huc / 1.0;
        qc[2 * G::NP2 + k1] = hvc / 1.0;
    }
}

// fc2d_geoclaw_fort_interpolate_face / _interpolate_corner
// (fc2d_geoclaw_interpolate_fort.f:24-403), sea_level = 0: the fine
// neighbour's ghost cells that the coarse boundary cells cover (a face: coarse
// cells 1..MX along it whose four fine cells lie in the fine array, which
// includes one hanging 2x2 corner; a corner: the corner cell), depth clipped
// at 0.  Coarse cells outside the coarse interior are read through
// ghost_source, as the halo pull reads them.
template <int MX, int MY>
__global__ void __launch_bounds__(64) k_interpolate_ghost(RegridArgs a)
{
    const int lo = a.range ? a.range[0] : 0, hi = a.range ? a.range[1] : a.n;
    for (int l = lo + (int)blockIdx.x; l < hi; l += gridDim.x) {
        double* q = link_buffer(a, l);
        if (!q) continue;
        RegridArgs b = a;  // the per-cell helpers read the state through b.q
        b.q = q;
        interpolate_link<MX, MY>(b, l);
    }
}

template <int MX, int MY>
__device__ __forceinline__ void interpolate_link(const RegridArgs& a, int l)
{
    using G = Geo<MX, MY>;
    const int side = a.side[l], igrid = a.igrid[l];
    int ox, oy;
    half_origin<MX, MY>(side, igrid, ox, oy);
    const int cp = a.cs[l];
    const double* ac = a.aux + (size_t)cp * G::NP2;
    double* qf = a.q + (size_t)a.fs[l] * 3 * G::NP2;
    const double* af = a.aux + (size_t)a.fs[l] * G::NP2;
    const int ncell = side >= 4 ? 1 : MX;
    for (int c = threadIdx.x; c < ncell; c += 64) {
        int ic, jc;
        if (side >= 4) {
            ic = (side - 4) & 1 ? MX : 1;
            jc = (side - 4) & 2 ? MY : 1;
        } else if (side < 2) {
            ic = side == 0 ? 1 : MX;
            jc = c + 1;
        } else {
            ic = c + 1;
            jc = side == 2 ? 1 : MY;
        }
        const int i0 = 2 * (ic - ox) - 1, j0 = 2 * (jc - oy) - 1;
        // fclaw2d_clawpatch_is_valid_interp: the four cells inside the fine array
        if (side < 4 && !(i0 >= 1 - MBC && i0 + 1 <= MX + MBC && j0 >= 1 - MBC && j0 + 1 <= MY + MBC)) continue;
        double qc[3], gx[3], gy[3];
        reconstruct_coarse<MX, MY>(a, cp, ac, ic, jc, qc, gx, gy);
#pragma unroll
        for (int mq = 0; mq < 3; mq++)
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const int ii = m & 1, jj = m >> 1;
                const double shiftx = (ii - 0.5) / 2.0, shifty = (jj - 0.5) / 2.0;
                const int k = G::ti(i0 + ii, j0 + jj);
                const double v = qc[mq] + gx[mq] * shiftx + gy[mq] * shifty;
                if (mq == 0) {
                    const double value = v - af[k];
                    qf[k] = value > 0.0 ? value : 0.0;
                } else {
                    qf[mq * G::NP2 + k] = v;
                }
            }
    }
}

// fc2d_geoclaw_fort_timeinterp (fc2d_geoclaw_timeinterp_fort.f:1-69): the
// two interior layers next to each face (every interior cell with i <= 2,
// i >= mx-1, j <= 2 or j >= my-1).
template <int MX, int MY>
__global__ void __launch_bounds__(256) k_timeinterp(TimeInterpArgs a)
{
    using G = Geo<MX, MY>;
    const int b = blockIdx.x;
    if (b >= a.n) return;
    const double* ql = a.qlast + (size_t)a.src[b] * 3 * G::NP2;
    const double* qc = a.qcurr + (size_t)a.src[b] * 3 * G::NP2;
    double* qd = a.qdst + (size_t)a.dst[b] * 3 * G::NP2;
    for (int c = threadIdx.x; c < MX * MY; c += 256) {
        const int i = c % MX + 1, j = c / MX + 1;
        if (!(i <= 2 || i >= MX - 1 || j <= 2 || j >= MY - 1)) continue;
        const int k = G::ti(i, j);
#pragma unroll
        for (int m = 0; m < 3; m++)
            qd[m * G::NP2 + k] = ql[m * G::NP2 + k] + a.alpha * (qc[m * G::NP2 + k] - ql[m * G::NP2 + k]);
    }
}

template <int MX, int MY>
__global__ void __launch_bounds__(256) k_hmax_list(const double* q, double* hmax, int* dry_age, const int* slots, int n)
{
    using G = Geo<MX, MY>;
    __shared__ double red[4];
    const int b = blockIdx.x;
    if (b >= n) return;
    const int s = slots[b];
    const double* src = q + (size_t)s * 3 * G::NP2;
    double m = 0.0;
    for (int e = threadIdx.x; e < MX * MY; e += 256) m = fmaxf64(m, src[G::ti(e % MX + 1, e / MX + 1)]);
    m = block_max<256>(m, red, threadIdx.x);
    if (threadIdx.x == 0) {
        hmax[s] = m;
        dry_age[s] = 0;
    }
}

// ---------------------------------------------------------------------------
// Pack / unpack between the reference's host layout (AoS q(m,i,j) with ghost
// cells, data_swap.f90:7-54) and the pool's SoA layout, one workgroup per
// patch.  ``stage`` holds n consecutive AoS patches.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_pack(const double* stage_q, const double* stage_aux,
                                              double* q, double* q2, double* aux, int first, int n,
                                              int np2, int maux)
{
    const int b = blockIdx.x;
    if (b >= n) return;
    const double* sq = stage_q + (size_t)b * 3 * np2;
    double* dq = q + (size_t)(first + b) * 3 * np2;
    double* dq2 = q2 ? q2 + (size_t)(first + b) * 3 * np2 : nullptr;
    for (int e = threadIdx.x; e < 3 * np2; e += blockDim.x) {
        const int c = e / 3, m = e - 3 * c;
        const double v = sq[e];
        dq[(size_t)m * np2 + c] = v;
        if (dq2) dq2[(size_t)m * np2 + c] = v;
    }
    if (stage_aux) {
        const double* sa = stage_aux + (size_t)b * maux * np2;
        double* da = aux + (size_t)(first + b) * np2;
        for (int c = threadIdx.x; c < np2; c += blockDim.x) da[c] = sa[(size_t)c * maux];
    }
}

// State of listed slots (q, ghost cells included) from one pool buffer to the
// other, one workgroup per slot.
__global__ void __launch_bounds__(256) k_copy_slots(const double* __restrict__ src, double* __restrict__ dst,
                                                    const int* __restrict__ slots, int n, int np2)
{
    const int b = blockIdx.x;
    if (b >= n) return;
    const size_t base = (size_t)slots[b] * 3 * np2;
    for (int e = threadIdx.x; e < 3 * np2; e += blockDim.x) dst[base + e] = src[base + e];
}

// State (q of one buffer, ghost cells included) and aux of slot src[b] to
// slot dst[b], one workgroup per slot (the destinations are distinct and no
// destination is a source, so the order does not matter).
__global__ void __launch_bounds__(256) k_move_slots(double* __restrict__ q, double* __restrict__ aux,
                                                    const int* __restrict__ src, const int* __restrict__ dst,
                                                    int n, int np2)
{
    const int b = blockIdx.x;
    if (b >= n) return;
    const size_t s = (size_t)src[b], d = (size_t)dst[b];
    for (int e = threadIdx.x; e < 3 * np2; e += blockDim.x) q[d * 3 * np2 + e] = q[s * 3 * np2 + e];
    for (int e = threadIdx.x; e < np2; e += blockDim.x) aux[d * np2 + e] = aux[s * np2 + e];
}

// The scenario of slots [first, first + n): the one field of their
// descriptors that re-targeting a scenario changes.
__global__ void __launch_bounds__(256) k_set_scen(PatchDesc* __restrict__ desc, int first, int n, int scen)
{
    const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k < n) desc[first + k].scen = scen;
}

__global__ void __launch_bounds__(256) k_unpack(double* stage_q, const double* q, int first,
                                                int n, int np2)
{
    const int b = blockIdx.x;
    if (b >= n) return;
    double* sq = stage_q + (size_t)b * 3 * np2;
    const double* dq = q + (size_t)(first + b) * 3 * np2;
    for (int e = threadIdx.x; e < 3 * np2; e += blockDim.x) {
        const int c = e / 3, m = e - 3 * c;
        sq[e] = dq[(size_t)m * np2 + c];
    }
}

// Maximum of n values into out[0] (one workgroup; batch CFL of the drop-in
// entry point).
__global__ void __launch_bounds__(256) k_max(const double* v, int n, double* out)
{
    __shared__ double red[4];
    double m = 0.0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) m = max_nan(m, v[k]);
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int k = 1; k < 4; k++) r = max_nan(r, red[k]);
        out[0] = r;
    }
}

// ---------------------------------------------------------------------------
// Host-side launchers (declared in gf_launch.h)
// ---------------------------------------------------------------------------
bool supported_shape(int mx, int my) { return mx == my && (mx == 8 || mx == 16 || mx == 32 || mx == 64); }

bool default_method(const Params& P)
{
    return P.order == 2 && P.transverse == 2 && P.mthlim[0] == 4 && P.mthlim[1] == 4 && P.mthlim[2] == 4;
}

// Step variants per patch shape: X(MX, TX, TY, NT, NPP, FREG) = tiles of
// TX x TY cells, NT threads per tile, NPP patches per workgroup (one-tile
// variants), x-face increments in registers (FREG 1) or LDS (0).  The first
// variant of a shape is its default; the autotuner picks among them by
// timing.  Every variant gives the same results bit for bit.
#define GF_VARIANTS(X)       \
    X(8, 8, 8, 128, 1, 0)       \
    X(8, 8, 8, 128, 1, 1)       \
    X(8, 8, 8, 128, 2, 1)       \
    X(8, 8, 8, 128, 3, 1)       \
    X(16, 16, 16, 384, 1, 0)    \
    X(16, 16, 16, 192, 1, 0)    \
    X(16, 16, 16, 384, 1, 1)    \
    X(16, 16, 16, 192, 1, 1)    \
    X(16, 16, 16, 384, 2, 1)    \
    X(32, 32, 32, 640, 1, 0)    \
    X(32, 16, 16, 384, 1, 0)    \
    X(32, 32, 32, 768, 1, 0)    \
    X(32, 32, 32, 704, 1, 0)    \
    X(32, 32, 32, 1024, 1, 0)   \
    X(32, 16, 16, 384, 1, 1)    \
    X(64, 32, 32, 640, 1, 0)    \
    X(64, 16, 16, 384, 1, 0)    \
    X(64, 32, 32, 768, 1, 0)    \
    X(64, 32, 16, 704, 1, 1)

static const StepVariant kVariants[] = {
#define GF_ROW(MX_, TX_, TY_, NT_, NPP_, FREG_) {MX_, MX_, TX_, TY_, NT_, NPP_, FREG_},
    GF_VARIANTS(GF_ROW)
#undef GF_ROW
};

int num_variants(int mx, int my)
{
    int n = 0;
    for (const StepVariant& v : kVariants) n += (v.mx == mx && v.my == my);
    return n;
}

bool get_variant(int mx, int my, int k, StepVariant* out)
{
    for (const StepVariant& v : kVariants)
        if (v.mx == mx && v.my == my && k-- == 0) {
            *out = v;
            return true;
        }
    return false;
}

static bool same_variant(const StepVariant& v, int mx, int tx, int ty, int nt, int npp, int freg)
{
    return v.mx == mx && v.tx == tx && v.ty == ty && v.nt == nt && v.npp == npp && v.freg == freg;
}

template <typename K>
static int occupancy_of(K kernel, int nt)
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, nt, 0) != hipSuccess) return 1;
    return n > 0 ? n : 1;
}

int wet_occupancy(int mx, int my, int variant)
{
    StepVariant v;
    if (mx != my || !get_variant(mx, my, variant, &v)) return 1;
#define GF_OCC(MX_, TX_, TY_, NT_, NPP_, FREG_)                                             \
    if (same_variant(v, MX_, TX_, TY_, NT_, NPP_, FREG_))                                  \
        return occupancy_of(k_all<MX_, MX_, TX_, TY_, NT_, NPP_, (bool)FREG_, true>, NT_ * NPP_);
    GF_VARIANTS(GF_OCC)
#undef GF_OCC
    return 1;
}

hipError_t launch_step(int mx, int my, int variant, const StepArgs& a, hipStream_t st)
{
    if (a.count <= 0) return hipSuccess;
    if (mx != my) return hipErrorInvalidValue;
    StepVariant v;
    if (!get_variant(mx, my, variant, &v)) return hipErrorInvalidValue;
    const int grid = a.count * (mx / v.tx) * (my / v.ty);
    const bool def = default_method(a.P);
    // the plain step runs one tile per workgroup: a variant's NPP does not apply
#define GF_STEP(MX_, TX_, TY_, NT_, NPP_, FREG_)                                             \
    if (same_variant(v, MX_, TX_, TY_, NT_, NPP_, FREG_)) {                                \
        if (def) k_step<MX_, MX_, TX_, TY_, NT_, (bool)FREG_, true><<<grid, NT_, 0, st>>>(a);  \
        else k_step<MX_, MX_, TX_, TY_, NT_, (bool)FREG_, false><<<grid, NT_, 0, st>>>(a);     \
        return hipGetLastError();                                                           \
    }
    GF_VARIANTS(GF_STEP)
#undef GF_STEP
    return hipErrorInvalidValue;
}

hipError_t launch_ghost(int mx, int my, const GhostArgs& a, hipStream_t st)
{
    if (a.count <= 0) return hipSuccess;
    switch (mx * 1000 + my) {
    case 32032: k_ghost<32, 32><<<a.count, 256, 0, st>>>(a); break;
    case 16016: k_ghost<16, 16><<<a.count, 256, 0, st>>>(a); break;
    case 8008: k_ghost<8, 8><<<a.count, 256, 0, st>>>(a); break;
    case 64064: k_ghost<64, 64><<<a.count, 256, 0, st>>>(a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_inflow(int mx, int my, const InflowArgs& a, int grid, hipStream_t st)
{
    if (grid <= 0) return hipSuccess;
    switch (mx * 1000 + my) {
    case 32032: k_inflow<32, 32><<<grid, 256, 0, st>>>(a); break;
    case 16016: k_inflow<16, 16><<<grid, 256, 0, st>>>(a); break;
    case 8008: k_inflow<8, 8><<<grid, 256, 0, st>>>(a); break;
    case 64064: k_inflow<64, 64><<<grid, 256, 0, st>>>(a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_classify(const BatchArgs* da, int slots, hipStream_t st)
{
    if (slots <= 0) return hipSuccess;
    k_classify<<<(slots + 255) / 256, 256, 0, st>>>(da);
    return hipGetLastError();
}

hipError_t launch_all(int mx, int my, int variant, bool def, const BatchArgs* da, int grid, hipStream_t st)
{
    StepVariant v;
    if (mx != my || !get_variant(mx, my, variant, &v)) return hipErrorInvalidValue;
#define GF_ALL(MX_, TX_, TY_, NT_, NPP_, FREG_)                                                                 \
    if (same_variant(v, MX_, TX_, TY_, NT_, NPP_, FREG_)) {                                                    \
        if (def) k_all<MX_, MX_, TX_, TY_, NT_, NPP_, (bool)FREG_, true><<<grid, NT_ * NPP_, 0, st>>>(da);     \
        else k_all<MX_, MX_, TX_, TY_, NT_, NPP_, (bool)FREG_, false><<<grid, NT_ * NPP_, 0, st>>>(da);        \
        return hipGetLastError();                                                                               \
    }
    GF_VARIANTS(GF_ALL)
#undef GF_ALL
    return hipErrorInvalidValue;
}


hipError_t launch_batch_control(const BatchArgs* da, int scenarios, hipStream_t st)
{
    k_batch_control<<<(scenarios + 63) / 64 > 0 ? (scenarios + 63) / 64 : 1, 64, 0, st>>>(da);
    return hipGetLastError();
}

#define GF_SIZES(CALL)                             \
    switch (mx * 1000 + my) {                      \
    case 32032: CALL(32, 32); break;               \
    case 16016: CALL(16, 16); break;               \
    case 8008: CALL(8, 8); break;                  \
    case 64064: CALL(64, 64); break;               \
    default: return hipErrorInvalidValue;          \
    }

hipError_t launch_average2coarse(int mx, int my, const RegridArgs& a, hipStream_t st)
{
    if (a.n <= 0) return hipSuccess;
#define GF_AVG(M, N) k_average2coarse<M, N><<<4 * a.n, 256, 0, st>>>(a)
    GF_SIZES(GF_AVG)
#undef GF_AVG
    return hipGetLastError();
}

hipError_t launch_interpolate2fine(int mx, int my, const RegridArgs& a, hipStream_t st)
{
    if (a.n <= 0) return hipSuccess;
#define GF_INT(M, N) k_interpolate2fine<M, N><<<4 * a.n, 256, 0, st>>>(a)
    GF_SIZES(GF_INT)
#undef GF_INT
    return hipGetLastError();
}

hipError_t launch_copy_slots(const double* src, double* dst, const int* slots, int n, int np2, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    k_copy_slots<<<n, 256, 0, st>>>(src, dst, slots, n, np2);
    return hipGetLastError();
}

hipError_t launch_average_ghost(int mx, int my, const RegridArgs& a, hipStream_t st)
{
    if (a.n <= 0) return hipSuccess;
#define GF_AG(M, N) k_average_ghost<M, N><<<a.n, 64, 0, st>>>(a)
    GF_SIZES(GF_AG)
#undef GF_AG
    return hipGetLastError();
}

hipError_t launch_interpolate_ghost(int mx, int my, const RegridArgs& a, hipStream_t st)
{
    if (a.n <= 0) return hipSuccess;
#define GF_IG(M, N) k_interpolate_ghost<M, N><<<a.n, 64, 0, st>>>(a)
    GF_SIZES(GF_IG)
#undef GF_IG
    return hipGetLastError();
}

hipError_t launch_timeinterp(int mx, int my, const TimeInterpArgs& a, hipStream_t st)
{
    if (a.n <= 0) return hipSuccess;
#define GF_TI(M, N) k_timeinterp<M, N><<<a.n, 256, 0, st>>>(a)
    GF_SIZES(GF_TI)
#undef GF_TI
    return hipGetLastError();
}

hipError_t launch_hmax_list(int mx, int my, const double* q, double* hmax, int* dry_age, const int* slots, int n,
                            hipStream_t st)
{
    if (n <= 0) return hipSuccess;
#define GF_HL(M, N) k_hmax_list<M, N><<<n, 256, 0, st>>>(q, hmax, dry_age, slots, n)
    GF_SIZES(GF_HL)
#undef GF_HL
    return hipGetLastError();
}
#undef GF_SIZES

hipError_t launch_frame_pack(int mx, int my, double* q, const double* aux, int first, int n, double* out,
                             hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    switch (mx * 1000 + my) {
    case 32032: k_frame_pack<32, 32><<<n, 256, 0, st>>>(q, aux, first, n, out); break;
    case 16016: k_frame_pack<16, 16><<<n, 256, 0, st>>>(q, aux, first, n, out); break;
    case 8008: k_frame_pack<8, 8><<<n, 256, 0, st>>>(q, aux, first, n, out); break;
    case 64064: k_frame_pack<64, 64><<<n, 256, 0, st>>>(q, aux, first, n, out); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_gauges(int mx, int my, const GaugeArgs& a, hipStream_t st)
{
    if (a.n <= 0) return hipSuccess;
    const int grid = (a.n + 63) / 64;
    switch (mx * 1000 + my) {
    case 32032: k_gauges<32, 32><<<grid, 64, 0, st>>>(a); break;
    case 16016: k_gauges<16, 16><<<grid, 64, 0, st>>>(a); break;
    case 8008: k_gauges<8, 8><<<grid, 64, 0, st>>>(a); break;
    case 64064: k_gauges<64, 64><<<grid, 64, 0, st>>>(a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_hmax(int mx, int my, const double* q, double* hmax, int first, int n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    switch (mx * 1000 + my) {
    case 32032: k_hmax<32, 32><<<n, 256, 0, st>>>(q, hmax, first, n); break;
    case 16016: k_hmax<16, 16><<<n, 256, 0, st>>>(q, hmax, first, n); break;
    case 8008: k_hmax<8, 8><<<n, 256, 0, st>>>(q, hmax, first, n); break;
    case 64064: k_hmax<64, 64><<<n, 256, 0, st>>>(q, hmax, first, n); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_control(ScenCtrl* ctrl, int nscen, const double* cfl, hipStream_t st)
{
    if (nscen <= 0) return hipSuccess;
    k_control<<<nscen, 256, 0, st>>>(ctrl, nscen, cfl);
    return hipGetLastError();
}

hipError_t launch_pack(const double* stage_q, const double* stage_aux, double* q, double* q2, double* aux,
                       int first, int n, int np2, int maux, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    k_pack<<<n, 256, 0, st>>>(stage_q, stage_aux, q, q2, aux, first, n, np2, maux);
    return hipGetLastError();
}

// Descriptors of host patches whose ghost cells come with them: cell sizes
// only, every side stored.
__global__ void __launch_bounds__(256) k_desc_fill(PatchDesc* d, const double* dx, const double* dy, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    PatchDesc p{};
    p.dx = dx[k];
    p.dy = dy[k];
    p.scen = 0;
    for (int s = 0; s < 8; s++) p.nbr[s] = NB_STORED;
    d[k] = p;
}

// The step's two powers (gf_pow.h) for n values: which 0 -> x^(7/3), 1 -> x^(2/3).
__global__ void __launch_bounds__(256) k_eval_powers(int which, const double* x, double* out, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = which == 0 ? pow_7_3(x[k]) : pow_2_3(x[k]);
}

hipError_t launch_eval_powers(int which, const double* x, double* out, int n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    k_eval_powers<<<(n + 255) / 256, 256, 0, st>>>(which, x, out, n);
    return hipGetLastError();
}

__global__ void __launch_bounds__(256) k_zero_words(int* p, size_t n)
{
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) p[k] = 0;
}

hipError_t launch_desc_fill(PatchDesc* d, const double* dx, const double* dy, int n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    k_desc_fill<<<(n + 255) / 256, 256, 0, st>>>(d, dx, dy, n);
    return hipGetLastError();
}

hipError_t launch_zero_words(int* p, size_t n, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    const size_t grid = std::min<size_t>((n + 255) / 256, 1024);
    k_zero_words<<<(unsigned)grid, 256, 0, st>>>(p, n);
    return hipGetLastError();
}

hipError_t launch_move_slots(double* q, double* aux, const int* src, const int* dst, int n, int np2, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    k_move_slots<<<n, 256, 0, st>>>(q, aux, src, dst, n, np2);
    return hipGetLastError();
}

hipError_t launch_set_scen(PatchDesc* desc, int first, int n, int scen, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    k_set_scen<<<(n + 255) / 256, 256, 0, st>>>(desc, first, n, scen);
    return hipGetLastError();
}

hipError_t launch_unpack(double* stage_q, const double* q, int first, int n, int np2, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    k_unpack<<<n, 256, 0, st>>>(stage_q, q, first, n, np2);
    return hipGetLastError();
}

hipError_t launch_max(const double* v, int n, double* out, hipStream_t st)
{
    k_max<<<1, 256, 0, st>>>(v, n, out);
    return hipGetLastError();
}

}  // namespace gf
