// gf_forest.cpp -- the quadtree of the resident-AMR driver, on the host.
//
// In GeoFlood, ForestClaw's p4est forest owns the leaves: fclaw2d_regrid
// (fc2d_geoclaw_run.c:569-575) refines tagged leaves, merges families of
// four siblings that all ask to coarsen, and closes the result under 2:1
// balance across faces and corners; the solver plugin then copies,
// interpolates and averages patches (fc2d_cpucuda.cpp:513-694).  Runs
// without ForestClaw (geoflood_amd.amr) keep this stand-in instead.
//
// A forest over an nx0 x ny0 brick of level-0 patches with levels
// 0..M is stored as the level of the covering leaf of every patch of the
// finest lattice: lev[J * nx + I], nx = nx0 2^M, ny = ny0 2^M.  A leaf
// (L, i, j) covers the finest patches [i s, (i+1) s) x [j s, (j+1) s),
// s = 2^(M - L).  Leaves are numbered in the order (L, j, i).
//
// Every function is a plain loop over the lattice (a few passes per
// regrid); arrays are int32, flags one byte per leaf, and adapt's work
// lattices one byte per patch.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

// left, right, bottom, top, left-bottom, right-bottom, left-top, right-top
constexpr int kDx[8] = {-1, 1, 0, 0, -1, 1, -1, 1};
constexpr int kDy[8] = {0, 0, -1, 1, -1, -1, 1, 1};

// The coarsest tree whose every finest patch has a level >= req there: each
// finest patch takes the smallest L whose level-L block around it has
// max(req) <= L (true at L = M when req <= M).  mx[k] holds the block maxima
// at the resolution of level M - k (mx[0] = req); out is the padded lattice
// (a border of -1 around nx x ny).  Levels are bytes (0..15).
struct Tree {
    int nx, ny, M;
    std::vector<std::vector<int8_t>> mx;
    std::vector<int8_t> out;   // (ny + 2) x (nx + 2)

    Tree(int nx_, int ny_, int M_) : nx(nx_), ny(ny_), M(M_), mx(M_ + 1), out((size_t)(nx_ + 2) * (ny_ + 2), -1)
    {
        for (int k = 0; k <= M; k++) mx[k].resize((size_t)(nx >> k) * (ny >> k));
    }

    // lv[k]: the level assigned to each block at the resolution of level
    // M - k, or -1 while none is (a block takes its parent's level, else L
    // when its maxima allow it)
    std::vector<std::vector<int8_t>> lv;

    void build(const std::vector<int8_t>& req)
    {
        std::copy(req.begin(), req.end(), mx[0].begin());
        for (int k = 1; k <= M; k++) {
            const int w = nx >> k, h = ny >> k, wp = nx >> (k - 1);
            const int8_t* a = mx[k - 1].data();
            int8_t* b = mx[k].data();
            for (int y = 0; y < h; y++) {
                const int8_t* __restrict__ r0 = a + (size_t)(2 * y) * wp;
                const int8_t* __restrict__ r1 = r0 + wp;
                int8_t* __restrict__ o = b + (size_t)y * w;
                for (int x = 0; x < w; x++)
                    o[x] = std::max(std::max(r0[2 * x], r0[2 * x + 1]), std::max(r1[2 * x], r1[2 * x + 1]));
            }
        }
        if (lv.empty()) {
            lv.resize(M + 1);
            for (int k = 0; k <= M; k++) lv[k].resize((size_t)(nx >> k) * (ny >> k));
        }
        {
            const int8_t* m0 = mx[M].data();
            int8_t* l0 = lv[M].data();
            for (size_t b = 0; b < lv[M].size(); b++) l0[b] = m0[b] <= 0 ? 0 : -1;
        }
        for (int L = 1; L <= M; L++) {
            const int k = M - L, w = nx >> k, h = ny >> k, wp = nx >> (k + 1);
            const int8_t* par = lv[k + 1].data();
            const int8_t* m = mx[k].data();
            int8_t* o = lv[k].data();
            for (int y = 0; y < h; y++) {
                const int8_t* __restrict__ pr = par + (size_t)(y >> 1) * wp;
                const int8_t* __restrict__ mr = m + (size_t)y * w;
                int8_t* __restrict__ orow = o + (size_t)y * w;
                for (int x = 0; x < w; x++) {
                    const int8_t pv = pr[x >> 1];
                    orow[x] = pv >= 0 ? pv : (mr[x] <= L ? (int8_t)L : (int8_t)-1);
                }
            }
        }
        const int W = nx + 2;
        for (int y = 0; y < ny; y++) std::memcpy(&out[(size_t)(y + 1) * W + 1], &lv[0][(size_t)y * nx], (size_t)nx);
    }
};

}  // namespace

extern "C" {

// The leaves of lev in the order (L, j, i): their level, block indices and
// origins on the finest lattice, and leaf_at[J * nx + I], the leaf covering
// each finest patch.  Returns the number of leaves, or -1 when lev is not a
// forest (a level outside [0, M], or a block not of one level).
int gff_index(int nx, int ny, int M, const int* lev, int* L, int* I, int* J, int* OI, int* OJ, int* leaf_at)
{
    if (M < 0 || M > 15 || nx <= 0 || ny <= 0 || nx % (1 << M) || ny % (1 << M)) return -1;
    const size_t cells = (size_t)nx * ny;
    for (size_t c = 0; c < cells; c++) {
        if (lev[c] < 0 || lev[c] > M) return -1;
        leaf_at[c] = -1;
    }
    int n = 0;
    for (int l = 0; l <= M; l++) {
        const int s = 1 << (M - l);
        for (int bj = 0; bj < ny / s; bj++)
            for (int bi = 0; bi < nx / s; bi++) {
                const int oi = bi * s, oj = bj * s;
                if (lev[(size_t)oj * nx + oi] != l) continue;
                for (int y = oj; y < oj + s; y++)
                    for (int x = oi; x < oi + s; x++) {
                        const size_t c = (size_t)y * nx + x;
                        if (lev[c] != l || leaf_at[c] != -1) return -1;
                        leaf_at[c] = n;
                    }
                L[n] = l;
                I[n] = bi;
                J[n] = bj;
                OI[n] = oi;
                OJ[n] = oj;
                n++;
            }
    }
    for (size_t c = 0; c < cells; c++)
        if (leaf_at[c] < 0) return -1;
    return n;
}

// Neighbour table and coarse/fine links of the n leaves of a forest (as
// gff_index returns them).  nbr[8 k + d] for the directions left, right,
// bottom, top, left-bottom, right-bottom, left-top, right-top: the
// same-level leaf there; else `stored` when the region lies inside the
// domain (covered by a coarser leaf or refined further); else the side's
// kind kinds[d] for a face and `wall` for a corner.  Links (coarse leaf,
// fine leaf, direction from the coarse one, igrid): every finer leaf across
// a side or corner of a coarser one, by direction, then by the fine leaf's
// half along the side (igrid, 0 the lower coordinate; 0 at corners), then
// by coarse leaf.  The link arrays hold 12 n entries.  Returns the number of
// links, or -1 when a finer neighbour is not one level finer (not 2:1
// balanced), -2 when a work list cannot be allocated.
//
// One pass over the leaves fills nbr and lists, per direction, the leaves
// with a finer neighbour there (in leaf order); the links are then written
// from those lists alone.
int gff_tables(int nx, int ny, int M, const int* lev, const int* leaf_at, int n, const int* L, const int* OI,
               const int* OJ, const int* kinds, int stored, int wall, int* nbr, int* lc, int* lf, int* ls, int* lg)
try {
    std::vector<int> finer[8];
    for (int k = 0; k < n; k++) {
        const int l = L[k], s = 1 << (M - l);
        for (int d = 0; d < 8; d++) {
            const int ni = OI[k] + kDx[d] * s, nj = OJ[k] + kDy[d] * s;
            int v;
            if (!(ni >= 0 && ni < nx && nj >= 0 && nj < ny)) {
                v = d < 4 ? kinds[d] : wall;
            } else {
                const size_t c = (size_t)nj * nx + ni;
                const int ln = lev[c];
                v = ln == l ? leaf_at[c] : stored;
                if (ln > l) finer[d].push_back(k);
            }
            nbr[8 * k + d] = v;
        }
    }
    int nl = 0;
    for (int d = 0; d < 8; d++) {
        if (finer[d].empty()) continue;
        const int dx = kDx[d], dy = kDy[d];
        // the fine leaves' halves next to the coarse leaf: both along the
        // side, the near one across it
        const int ci0 = dx < 0 ? 1 : 0, ci1 = dx > 0 ? 0 : 1;
        const int cj0 = dy < 0 ? 1 : 0, cj1 = dy > 0 ? 0 : 1;
        for (int cj = cj0; cj <= cj1; cj++)
            for (int ci = ci0; ci <= ci1; ci++)
                for (const int k : finer[d]) {
                    const int s = 1 << (M - L[k]), h = s / 2;
                    const int ki = OI[k] + dx * s + ci * h, kj = OJ[k] + dy * s + cj * h;
                    const size_t c = (size_t)kj * nx + ki;
                    if (lev[c] != L[k] + 1) return -1;
                    lc[nl] = k;
                    lf[nl] = leaf_at[c];
                    ls[nl] = d;
                    lg[nl] = d < 4 ? (d < 2 ? cj : ci) : 0;
                    nl++;
                }
    }
    return nl;
} catch (...) {
    return -2;  // no exception crosses the C ABI (the work lists' allocation)
}

// A forest's neighbour table laid out in pool slots: leaf k sits in slot
// base + perm[k].  inv[r] receives the leaf in slot base + r, and row r of
// out that leaf's row of nbr with every leaf entry v >= 0 replaced by its
// slot base + perm[v] (negative codes kept).  Returns 0, or -1 when perm is
// not a permutation of [0, n) or an entry of nbr is not below n.
int gff_slot_table(int n, const int* nbr, const long long* perm, int base, long long* inv, int* out)
{
    for (int r = 0; r < n; r++) inv[r] = -1;
    for (int k = 0; k < n; k++) {
        const long long r = perm[k];
        if (r < 0 || r >= n || inv[r] >= 0) return -1;
        inv[r] = k;
    }
    for (int r = 0; r < n; r++) {
        const int* row = nbr + 8 * inv[r];
        int* o = out + 8 * (size_t)r;
        for (int d = 0; d < 8; d++) {
            const int v = row[d];
            if (v >= n) return -1;
            o[d] = v >= 0 ? base + (int)perm[v] : v;
        }
    }
    return 0;
}

// The levels after one regrid of the forest lev (its n leaves L, I, J, OI,
// OJ and leaf_at as gff_index returns them): leaves with refine[k] split
// once (below M); a family of four sibling leaves of level L > minlevel that
// all have coarsen[k] and none refine[k] merges into its parent; then the
// coarsest 2:1-balanced tree (faces and corners) at least as fine as that
// request, which may refine further leaves or keep a family from merging.
// Writes out[ny * nx]; returns 1 when it differs from lev, else 0 (-2 when
// its work arrays cannot be allocated).
int gff_adapt(int nx, int ny, int minlevel, int M, const int* lev, const int* leaf_at, int n, const int* L,
              const int* I, const int* J, const int* OI, const int* OJ, const unsigned char* refine,
              const unsigned char* coarsen, int* out)
try {
    const size_t cells = (size_t)nx * ny;
    std::vector<int8_t> req(cells);
    for (size_t c = 0; c < cells; c++) req[c] = (int8_t)lev[c];
    bool request = false;
    for (int k = 0; k < n; k++) {
        if (!refine[k] || L[k] >= M) continue;
        const int s = 1 << (M - L[k]);
        for (int y = OJ[k]; y < OJ[k] + s; y++) std::memset(&req[(size_t)y * nx + OI[k]], L[k] + 1, (size_t)s);
        request = true;
    }
    // families: the leaf at the lower-left of its parent, and its three
    // siblings (leaves of the same level at the parent's other quarters)
    for (int k = 0; k < n; k++) {
        if (L[k] <= minlevel || (I[k] & 1) || (J[k] & 1)) continue;
        const int s = 1 << (M - L[k]);
        bool all = true;
        for (int q = 0; q < 4 && all; q++) {
            const size_t c = (size_t)(OJ[k] + (q >> 1) * s) * nx + OI[k] + (q & 1) * s;
            const int m = leaf_at[c];
            all = lev[c] == L[k] && coarsen[m] && !refine[m];
        }
        if (!all) continue;
        for (int y = OJ[k]; y < OJ[k] + 2 * s; y++) std::memset(&req[(size_t)y * nx + OI[k]], L[k] - 1, (size_t)(2 * s));
        request = true;
    }
    if (!request) {
        // the request is the forest itself, a balanced tree: the closure keeps it
        std::memcpy(out, lev, sizeof(int) * cells);
        return 0;
    }
    // balance closure
    const int8_t lo = (int8_t)minlevel;
    for (size_t c = 0; c < cells; c++) req[c] = std::max(req[c], lo);
    Tree t(nx, ny, M);
    const int W = nx + 2;
    for (;;) {
        t.build(req);
        int ok = 1;
        for (int y = 0; y < ny; y++) {
            const int8_t* __restrict__ up = &t.out[(size_t)y * W + 1];
            const int8_t* __restrict__ mid = up + W;
            const int8_t* __restrict__ dn = mid + W;
            int8_t* __restrict__ rq = &req[(size_t)y * nx];
            for (int x = 0; x < nx; x++) {
                const int8_t nb = std::max(std::max(std::max(up[x - 1], up[x]), std::max(up[x + 1], mid[x - 1])),
                                           std::max(std::max(mid[x + 1], dn[x - 1]), std::max(dn[x], dn[x + 1])));
                const int8_t need = (int8_t)(nb - 1);
                ok &= mid[x] >= need;
                rq[x] = std::max(rq[x], need);
            }
        }
        if (ok) break;
    }
    int changed = 0;
    for (int y = 0; y < ny; y++) {
        const int8_t* row = &t.out[(size_t)(y + 1) * W + 1];
        int* o = out + (size_t)y * nx;
        const int* l0 = lev + (size_t)y * nx;
        for (int x = 0; x < nx; x++) {
            o[x] = row[x];
            changed |= o[x] != l0[x];
        }
    }
    return changed;
} catch (...) {
    return -2;  // no exception crosses the C ABI (the work vectors' allocation)
}

}  // extern "C"
