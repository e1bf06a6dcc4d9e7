// Mixed-precision solve (FP64 iterative refinement around an FP32 inner PCG
// with an FP32 node-gather or split-BF16 MFMA level-0 operator).
#pragma once

#include <vector>

#include "gputo_internal.h"
#include "vcycle.h"

namespace gputo {

struct MfmaState;

struct MixedState {
    HierView<float> H;                 // FP32 hierarchy view
    float *A32[kMaxLevels] = {};       // FP32 packed element matrices
    float *S32[kMaxLevels] = {};       // FP32 node stencils of the levels that have them
    float *CXl32[kMaxLevels] = {};     // FP32 element scales of matrix-free coarse levels
    float *Ke32 = nullptr, *CX32 = nullptr;
    float *X = nullptr, *R = nullptr, *P = nullptr, *Q = nullptr, *Z = nullptr;  // inner CG vectors
    float *Pcg = nullptr;              // the FP64 PCG's last kCgLog search directions, kept in FP32
    double *alog = nullptr;            // their alphas (0 for converged iterations)
    CgScalars *sc = nullptr, *sc_host = nullptr;
    double *norm = nullptr;            // ||r|| of the outer residual (device)
    double norm_host[2] = {0, 0};
    MfmaState *mf = nullptr;
    hipGraphExec_t vc_exec = nullptr;  // the recorded V-cycle of the current solve (mixed_vcycle_graph)
    int vc_kind = 0;                   // its learned stage (learn_stage_kind)
    const int *vc_ctl = nullptr;
    bool vc_pre0 = false;              // recorded without the level-0 prerelaxation pass
    int last_inner = 0, last_outer = 0, last_learned_ok = 0, last_learned_bad = 0;
    long long bytes = 0;
    std::vector<void *> allocs;
};

void mixed_create(gputo_handle *h);
void mixed_destroy(gputo_handle *h);
void mixed_stats(gputo_handle *h, gputo_to_stats *st);
void mixed_begin(gputo_handle *h);                                        // FP32 hierarchy copies
// z = V32(r); r32_ready: the FP32 copy of r is already in mixed_r32(h);
// z null: the result stays in FP32 (mixed_z32(h)) for consumers that widen it
// pre0_ready: the prerelaxed level-0 iterate is already in
// mixed_prerelax_target(h) (formed from mixed_r32(h) by the caller)
void mixed_vcycle(gputo_handle *h, const double *r, double *z, const int *ctl, bool r32_ready = false,
                  bool pre0_ready = false);
// Where the FP32 V-cycle's first level-0 pass writes the prerelaxed iterate
// prerelax*F/AD (its sweeps then start from it), with that AD and prerelax;
// null when the V-cycle does not start with that pass.
float *mixed_prerelax_target(gputo_handle *h, const float **ad, double *prerelax);
// The same V-cycle with r32_ready and z null (params.graph_vcycles), its
// launches recorded once in a HIP graph and replayed with one launch while
// they stay the same: the learned stage launches nothing, or the
// correction, in every V-cycle (a logged sample is launched from the host).
// Results equal mixed_vcycle's bit for bit.  The graph holds the
// hierarchy's pointers and values: it is dropped by mixed_begin, mixed_end
// and mixed_graph_drop.
void mixed_vcycle_graph(gputo_handle *h, const int *ctl, bool pre0_ready = false);
void mixed_graph_drop(gputo_handle *h);
float *mixed_r32(gputo_handle *h);        // the FP32 right-hand side of the V-cycle
const float *mixed_z32(gputo_handle *h);  // the FP32 result of the last V-cycle
float *mixed_p32(gputo_handle *h);        // the PCG's FP32 search directions (kCgLog slots of ndof)
double *mixed_alog(gputo_handle *h);      // and their alphas
void mixed_end(gputo_handle *h, int nit, int converged);
int mixed_apply_l0(gputo_handle *h, int kind, const double *u, double *y);
int mixed_bench_operator(gputo_handle *h, int kind, int reps, double *avg_ms);
HierView<float> mixed_level_view(gputo_handle *h);

// split-BF16 MFMA level-0 operator (mfma.hip)
void mfma_create(gputo_handle *h, MixedState *m);
void mfma_destroy(gputo_handle *h, MixedState *m);
int mfma_blocks(const gputo_handle *h);             // blocks (= partials) of one launch
long mfma_tile_blocks(const Grid &g0);              // the same, before the state exists

}  // namespace gputo
