// gf_plugin.cpp -- host-side mirror of the fc2d_cpu_cuda plugin's GPU route.
//
// ForestClaw iterates over the patches of a level and calls the plugin's
// single_step_update slot once per patch with buffer_data = {iter,
// total_count} (fc2d_cpucuda.cpp:409-495).  The reference buffers patch
// descriptors (cudaclaw_store_buffer, cudaclaw_store_patches.cu:13-42) and
// flushes a batch through cudaclaw_step2_batch when buffer_len patches are
// queued or the last patch arrives; non-flushing calls return a CFL of 0 and
// the caller takes the maximum over all calls.  This file keeps that
// protocol (same argument meaning and return values) on top of
// gf_step2_batch, with two deliberate differences, both following the
// Fortran route rather than the CUDA one: the MAXCFL_CAP (1000 -> 0) is
// applied per patch (fc2d_cpucuda.cpp:392) instead of per batch (:492), and
// errors are returned as status codes instead of exit().
#include <cstring>
#include <new>
#include <vector>

#include "../../include/geoflood_amd.h"

struct gf_plugin {
    int mx, my, mbc, maux, buffer_len;
    std::vector<double*> q, aux;  // queued host patches (the reference's flux_array)
    std::vector<double> dx, dy;
};

extern "C" {

gf_plugin* gf_plugin_create(int mx, int my, int mbc, int maux, int buffer_len)
{
    if (mx <= 0 || my <= 0 || mbc != 2 || maux < 1 || buffer_len <= 0) return nullptr;
    gf_plugin* p = new (std::nothrow) gf_plugin();
    if (!p) return nullptr;
    p->mx = mx;
    p->my = my;
    p->mbc = mbc;
    p->maux = maux;
    p->buffer_len = buffer_len;
    try {  // no exception crosses the C ABI
        p->q.reserve(buffer_len);
        p->aux.reserve(buffer_len);
        p->dx.reserve(buffer_len);
        p->dy.reserve(buffer_len);
    } catch (...) {
        delete p;
        return nullptr;
    }
    return p;
}

int gf_plugin_destroy(gf_plugin* p)
{
    delete p;
    return GF_OK;
}

int gf_plugin_single_step_update(gf_plugin* p, double* q, double* aux, double dx, double dy,
                                 int iter, int total_count, double t, double dt, double* maxcfl)
try {
    if (!p || !q || !aux || !maxcfl) return GF_EINVAL;
    if (iter < 0 || total_count <= 0 || iter >= total_count) return GF_EINVAL;
    *maxcfl = 0.0;
    if (iter == 0) {  // a new sweep over the level: the buffer starts empty
        p->q.clear();
        p->aux.clear();
        p->dx.clear();
        p->dy.clear();
    }
    p->q.push_back(q);
    p->aux.push_back(aux);
    p->dx.push_back(dx);
    p->dy.push_back(dy);
    const bool full = ((iter + 1) % p->buffer_len) == 0;
    const bool last = (iter + 1) == total_count;
    if (!full && !last) return GF_OK;
    const int n = (int)p->q.size();
    int rc = gf_step2_batch(p->mx, p->my, p->mbc, p->maux, n, p->q.data(), p->aux.data(),
                            p->dx.data(), p->dy.data(), t, dt, maxcfl);
    p->q.clear();
    p->aux.clear();
    p->dx.clear();
    p->dy.clear();
    return rc;
} catch (...) {  // a queue that cannot grow: no exception crosses the C ABI
    return GF_ENOMEM;
}

}  // extern "C"
