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
};

struct GhostArgs {
    const PatchDesc* desc;
    int first, count;
    double* q0;
    double* q1;
    ScenCtrl* ctrl;  // when set, each patch uses its scenario's current buffer
    int cur;         // otherwise this buffer
};

bool supported_shape(int mx, int my);
hipError_t launch_step(int mx, int my, const StepArgs& a, hipStream_t st);
hipError_t launch_ghost(int mx, int my, const GhostArgs& a, hipStream_t st);
hipError_t launch_control(ScenCtrl* ctrl, int nscen, const double* cfl, hipStream_t st);
hipError_t launch_pack(const double* stage_q, const double* stage_aux, double* q, double* aux,
                       int first, int n, int np2, int maux, hipStream_t st);
hipError_t launch_unpack(double* stage_q, const double* q, int first, int n, int np2, hipStream_t st);
hipError_t launch_max(const double* v, int n, double* out, hipStream_t st);

}  // namespace gf
