// gf_ghost.h -- patch geometry in LDS/HBM and where a patch's ghost cells
// come from on the device.
#pragma once

#include "gf_kernels.h"

namespace gf {

// Geometry of an MX x MY patch (or tile) with its 2-cell halo.
template <int MX, int MY>
struct Geo {
    static constexpr int NXP = MX + 2 * MBC;
    static constexpr int NYP = MY + 2 * MBC;
    static constexpr int NP2 = NXP * NYP;
    static constexpr int PX = MX + 3;
    static constexpr int PY = MY + 3;
    static constexpr int NPOS = PX * PY;
    // x-face flux increments kept in LDS between the x and y phases:
    // faces I in [1, MX+1] of rows j in [1, MY]
    static constexpr int NF = (MX + 1) * MY;
    __device__ static constexpr int fi_(int I, int j) { return (j - 1) * (MX + 1) + (I - 1); }
    // LDS index of Fortran cell (fi, fj), fi in [-1, MX+2]
    __device__ static constexpr int ti(int fi, int fj) { return (fj + 1) * NXP + (fi + 1); }
};

// Source of ghost cell (fi, fj) of a patch (Fortran indices, outside the
// interior): the pool index and cell whose value ForestClaw's same-level face
// copy, corner copy and bc2 passes (left, right, bottom, top over full
// extents; geoflood_bc2_fort.f90:118-131 and twins) leave there, and whether
// hu / hv change sign.  The cell is another patch's interior cell, or, for a
// corner between a physical side and a stored side, the patch's own stored
// ghost cell that bc2 mirrors (the stored strips are filled first: the
// coarse/fine transfers run before the physical boundaries, as in
// ForestClaw's ghost update).  Returns false when the ghost is read from the
// patch's own ghost cells: supplied by the host or a coarse/fine transfer
// (NB_STORED), written by k_inflow (an NB_INFLOW side or a corner next to
// one), or a corner with no rule (two stored sides).
template <int MX, int MY>
__device__ __forceinline__ bool ghost_source(const int* __restrict__ nbr, int patch, int fi, int fj,
                                             int& sp, int& si, int& sj, bool& neg_u, bool& neg_v)
{
    const int ox = fi < 1 ? -1 : (fi > MX ? 1 : 0);
    const int oy = fj < 1 ? -1 : (fj > MY ? 1 : 0);
    const int nx = ox < 0 ? nbr[0] : (ox > 0 ? nbr[1] : 0);
    const int ny = oy < 0 ? nbr[2] : (oy > 0 ? nbr[3] : 0);
    const bool px = (nx == NB_WALL || nx == NB_EXTRAP);
    const bool py = (ny == NB_WALL || ny == NB_EXTRAP);
    const int txv = (nx == NB_WALL) ? (ox < 0 ? 1 - fi : 2 * MX + 1 - fi) : (ox < 0 ? 1 : MX);
    const int tyv = (ny == NB_WALL) ? (oy < 0 ? 1 - fj : 2 * MY + 1 - fj) : (oy < 0 ? 1 : MY);
    neg_u = false;
    neg_v = false;
    if (ox != 0 && oy != 0) {
        const int nc = oy < 0 ? (ox < 0 ? nbr[4] : nbr[5]) : (ox < 0 ? nbr[6] : nbr[7]);
        if (nx == NB_INFLOW || ny == NB_INFLOW) {
            return false;  // bc2's inflow side pass overwrites the corner; k_inflow stores it
        } else if (nc >= 0) {
            sp = nc; si = fi - ox * MX; sj = fj - oy * MY;
        } else if (px && py) {
            sp = patch; si = txv; sj = tyv;
            neg_u = (nx == NB_WALL); neg_v = (ny == NB_WALL);
        } else if (px && ny >= 0) {
            sp = ny; si = txv; sj = fj - oy * MY;
            neg_u = (nx == NB_WALL);
        } else if (py && nx >= 0) {
            sp = nx; si = fi - ox * MX; sj = tyv;
            neg_v = (ny == NB_WALL);
        } else if (px && ny == NB_STORED) {
            // bc2's x pass runs over the full extent after the stored strip
            // was filled (a coarse/fine transfer or the host): its mirror
            sp = patch; si = txv; sj = fj;
            neg_u = (nx == NB_WALL);
        } else if (py && nx == NB_STORED) {
            sp = patch; si = fi; sj = tyv;
            neg_v = (ny == NB_WALL);
        } else {
            return false;
        }
    } else if (ox != 0) {
        if (nx >= 0) { sp = nx; si = fi - ox * MX; sj = fj; }
        else if (px) { sp = patch; si = txv; sj = fj; neg_u = (nx == NB_WALL); }
        else return false;
    } else {
        if (ny >= 0) { sp = ny; si = fi; sj = fj - oy * MY; }
        else if (py) { sp = patch; si = fi; sj = tyv; neg_v = (ny == NB_WALL); }
        else return false;
    }
    return true;
}

}  // namespace gf
