// Learned level-2 correction of the FP32 V-cycle: a small 3-D convolutional
// encoder-decoder on the level-2 node grid, trained on the device with this
// library's own forward and backward kernels.
#pragma once

#include <utility>
#include <vector>

#include "gputo_internal.h"
#include "vcycle.h"

namespace gputo {

// Spatial extent of one activation: n0 x n1 x n2 = (i, k, j) node counts,
// position p = (i*n1 + k)*n2 + j (the level's node order).
struct Dims3 {
    int n0, n1, n2;
    __host__ __device__ long size() const { return (long)n0 * n1 * n2; }
};

// Layers of the network (3x3x3 kernels, padding 1):
//   e1 = relu(conv1(x))        4 -> 16, stride 1   (full grid)
//   e2 = relu(conv2(e1))      16 -> 32, stride 2
//   e3 = relu(conv3(e2))      32 -> 64, stride 2
//   d2 = relu(conv4(e3))      64 -> 32, stride 1
//   u2 = up(d2) + e2
//   d1 = relu(conv5(u2))      32 -> 16, stride 1
//   u1 = up(d1) + e1
//   y  = conv6(u1)            16 -> 3,  stride 1   (linear)
constexpr int kNetLayers = 6;
struct NetLayer {
    int cin, cout, stride, relu;
    long w_off, b_off;  // offsets into the flat parameter array
};

struct LearnState {
    Dims3 d0{}, d1{}, d2{};            // full, /2, /4 grids
    NetLayer L[kNetLayers];
    long nparam = 0;
    float *P = nullptr, *G = nullptr, *M = nullptr, *V = nullptr;  // params, grads, Adam moments
    int adam_t = 0;
    // activations (forward cache) and their gradients
    float *x0 = nullptr, *e1 = nullptr, *e2 = nullptr, *e3 = nullptr, *a2 = nullptr, *u2 = nullptr,
          *a1 = nullptr, *u1 = nullptr, *y = nullptr;
    float *gy = nullptr, *gu1 = nullptr, *ga1 = nullptr, *gu2 = nullptr, *ga2 = nullptr, *ge3 = nullptr,
          *ge2 = nullptr, *ge1 = nullptr;
    // level-2 DOF vectors of the correction stage
    float *d = nullptr, *e = nullptr, *Ae = nullptr;
    double *sc = nullptr;        // device scalars: sigma, gamma, loss, dots
    int *count = nullptr;        // device counters: accepted, rejected
    double *partials = nullptr;  // reduction partials
    float *wpart = nullptr;      // weight-gradient partials per position chunk
    // logged level-2 defects (training samples) of the current solve
    int nsamp_max = 0, nsamp = 0, vc_count = 0;
    float *samples = nullptr;
    bool logging = true;         // log the defects of the first nsamp_max V-cycles of a solve
    bool trained = false;
    // whether the current solve applies the stage, and the solve-level
    // fallback of params.learned = 2 (LearnPolicy in learn.hip)
    bool use = true;
    int pol_state = 0, pol_off_left = 0, pol_backoff = 0, pol_on_count = 0;
    double pol_t_plain = 0.0;
    // (seconds, relres) at the host's polls: of the last plain solve, and of
    // the current solve; a trial that falls behind the first is cut short
    std::vector<std::pair<double, double>> pol_ref, pol_cur;
    bool pol_abort = false;
    long long bytes = 0;
    std::vector<void *> allocs;
};

void learn_create(gputo_handle *h);
void learn_destroy(gputo_handle *h);
void learn_init_params(gputo_handle *h, unsigned seed);
void learn_policy_reset(LearnState *L);  // the fallback policy starts over (REF)
// inputs of the correction stage's forward pass: the feature kernel's
// partials of sum d^2/ad (norm_n of them, over norm_count DOFs), the
// diagonal and fixed-DOF mask of the level, and the correction's output
struct CorrIO {
    const double *norm;
    int norm_n;
    double norm_count;
    const float *ad;
    const uint8_t *mask;
    float *e;
};
void learn_forward(gputo_handle *h, const int *ctl, const CorrIO *io);  // y = N(x0), activations cached
void learn_backward(gputo_handle *h);                  // G += d<gy, y>/dtheta
// Extra stage after the level-2 post-smoothing of the FP32 V-cycle:
// d = F - A U;  e = N(d);  gamma = <d,e>/<e,Ae>;  U += gamma e if gamma > 0.
// Logs d as a training sample when logging is on.
void learn_stage(gputo_handle *h, const HierView<float> &H, int l, const float *F, float *U, const int *ctl);
// What learn_stage launches in the next V-cycle: -1 a logged sample (its
// slot differs from V-cycle to V-cycle), 0 nothing, 1 the correction.
int learn_stage_kind(const gputo_handle *h);
// A V-cycle whose launches were replayed: counted as learn_stage counts it.
void learn_stage_count(gputo_handle *h);
void learn_begin_solve(gputo_handle *h);                   // reset counters and the sample log; policy
// after a solve of `seconds` (wall): counters, logged samples, policy update
void learn_end_solve(gputo_handle *h, int *accepted, int *rejected, double seconds = -1.0);
// progress of the current solve at a host poll (CG iteration `it`, wall
// seconds since its start, relative residual)
void learn_poll(gputo_handle *h, int it, double seconds, double relres);
int learn_train(gputo_handle *h, int steps, double lr, double *loss_first, double *loss_last);

}  // namespace gputo
