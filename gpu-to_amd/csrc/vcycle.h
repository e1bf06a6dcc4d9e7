// The multigrid V-cycle, written once over a typed view of the hierarchy so
// the FP64 solve and the FP32 inner solve of the mixed-precision path share
// it (vcycle, Cantilever/TOP3D125MGCGCUDA/mgcg1.cu:468-651;
// Cantilever/omp/mgcg1.c:283-447).
#pragma once

#include "gputo_internal.h"

namespace gputo {

// Smoother / hierarchy constants of a preset.
struct Preset {
    int npre, npost;
    double omega, prerelax;
    int l1_galerkin, zero_after_interp, corr_mirrored;
    int mg9;           // every level matrix-free with coarse element scales
    double cx_floor;   // smoother defects skip elements with CX below this
};

inline Preset preset_of(const gputo_params &prm)
{
    // REF_C: mgcg1.c:110,308,353,363-364, mgcg1.cu:188,631,1196,1205;
    // MATLAB_CFG0: top3D125mgcg.m:147,194,243 (omega 0.6, one sweep from zero);
    // MGCG9: mgcg9.cu:920-937,939-996 (0.7 prerelax, omega 0.33, nswp sweeps)
    if (prm.preset == GPUTO_PRESET_MATLAB_CFG0) return {0, 1, 0.6, 0.6, 1, 1, 0, 0, 0.0};
    if (prm.preset == GPUTO_PRESET_MGCG9) {
        const int n = prm.nswp > 0 ? prm.nswp : 2;
        return {n, n, 0.33, 0.7, 0, 1, 0, 1, 1e-8};
    }
    return {3, 3, 0.48, 0.7, 0, 1, 1, 0, 0.0};
}

// Level-0 operator implementation of a view.
// (values match gputo_params.l0_kernel; MFMA flavour = kind - L0_MFMA_F32)
enum L0Kind { L0_AUTO = 0, L0_GATHER = 1, L0_MFMA_F32 = 2, L0_MFMA_BF16 = 3, L0_MFMA_BF16X3 = 4, L0_MFMA_BF16X1 = 5 };

template <class T>
struct LevelView {
    T *U = nullptr, *F = nullptr, *R = nullptr, *AD = nullptr, *Tmp = nullptr;
    const T *A = nullptr;   // packed element matrices (levels >= 1)
    const T *S = nullptr;   // node stencils of a stored level, or null
    const T *CX = nullptr;  // element scales of a matrix-free level (mgcg9)
};

template <class T>
struct HierView {
    LevelView<T> L[kMaxLevels];
    const T *Ke = nullptr;   // level-0 element matrix (24x24 row-major)
    const T *CX = nullptr;   // level-0 element scales
    int l0 = L0_GATHER;
    const void *l0_ctx = nullptr;  // state of a non-gather level-0 kernel
    bool l1_mf = false;            // level 1 applied matrix-free from CX
    bool learn = false;            // learned level-2 stage (FP32 views only)
    bool mg9 = false;              // every level matrix-free (mgcg9 engine)
    bool per = false;              // periodic cell: folded defects, images kept equal
    bool pre0_ready = false;       // level 0's prerelaxed iterate is already formed (by the caller)
    int nb = 1;                    // right-hand sides: level vectors hold nb vectors each, one after another
};

constexpr int kLearnLevel = 2;
// learned correction after the level-2 post-smoothing (learn.hip)
void learn_stage(gputo_handle *h, const HierView<float> &H, int l, const float *F, float *U, const int *ctl);

// level-0 BF16-MFMA operator (mfma.hip); ctx is the MixedState
void launch_op0_mfma(hipStream_t s, int mode, const OpArgsT<float> &a, const void *ctx, int flavor);

// matrix-free rediscretised level-1 operator (l1op.hip); CXf = level-0 CX
template <class T>
void launch_op1_mf(hipStream_t s, int mode, const OpArgsT<T> &a, const T *CXf, const Tile &t, double nu);
Tile l1_tile(const Grid &g);    // FP32 tiling
Tile l1_tile64(const Grid &g);  // FP64 tiling
constexpr double kL1Nu = 0.3;  // Poisson ratio of the level-1 rediscretisation (mgcg1.c:1539-1584)

// Whether level 1 runs matrix-free under the handle's parameters.
inline bool use_l1_mf(const gputo_params &prm, int nl)
{
    if (nl < 3) return false;  // level 1 is the coarsest: its stored matrix feeds the Cholesky
    if (preset_of(prm).mg9) return false;
    const bool rediscretised = !preset_of(prm).l1_galerkin;
    if (prm.l1_kernel == 1) return false;
    return rediscretised;
}

template <class T>
OpArgsT<T> view_args(const gputo_handle *h, const HierView<T> &H, int l, const int *ctl)
{
    OpArgsT<T> a{};
    a.g = h->L[l].g;
    a.t = h->L[l].t;
    a.Ke = H.Ke;
    a.CX = H.CX;
    a.A = H.L[l].A;
    a.S = H.L[l].S;
    a.mask = h->L[l].mask;
    a.AD = H.L[l].AD;
    a.ctl = ctl;
    a.nb = H.nb;
    const Preset p = preset_of(h->prm);
    a.omega = p.omega;
    a.prerelax = p.prerelax;
    if (p.mg9) {
        a.CX = H.L[l].CX;
        a.t = h->L[l].tg;
        a.cx_floor = p.cx_floor;
    }
    a.kc = h->ke_classes;
    for (int k = 0; k < kKeMaxClasses; ++k) a.kv[k] = (T)h->kv[k];
    return a;
}

template <class T>
void level_op(const gputo_handle *h, const HierView<T> &H, int l, int mode, const OpArgsT<T> &a)
{
    if (H.mg9) {
        launch_op0<T>(h->stream, mode, a);  // level scales and the CX floor are in a
    } else if (l == 1 && H.l1_mf) {
        launch_op1_mf<T>(h->stream, mode, a, H.CX, std::is_same<T, double>::value ? h->L[1].tmf64 : h->L[1].tmf,
                         kL1Nu);
    } else if (l > 0) {
        launch_opA<T>(h->stream, mode, a);
    } else if constexpr (std::is_same<T, float>::value) {
        if (H.l0 >= L0_MFMA_F32)
            launch_op0_mfma(h->stream, mode, a, H.l0_ctx, H.l0 - L0_MFMA_F32);
        else
            launch_op0<T>(h->stream, mode, a);
    } else {
        launch_op0<T>(h->stream, mode, a);
    }
}

// Periodic sweeps: each is the non-periodic defect into R, then the
// Jacobi update of every master from its folded defect, copied to its
// images; `pre` first sets U = prerelax*F/AD.
template <class T>
void sweeps_per(const gputo_handle *h, const HierView<T> &H, int l, const T *F, T *U, int n, bool pre, const int *ctl)
{
    const LevelView<T> &v = H.L[l];
    const Level &lv = h->L[l];
    const Preset p = preset_of(h->prm);
    if (pre) launch_per_jacobi<T>(h->stream, lv.g, U, v.R, F, v.AD, lv.mask, p.omega, p.prerelax, 1, ctl, H.nb);
    for (int s = 0; s < n; ++s) {
        OpArgsT<T> a = view_args(h, H, l, ctl);
        a.U = U;
        a.F = F;
        a.out = v.R;
        level_op(h, H, l, OP_DEFECT, a);
        launch_per_jacobi<T>(h->stream, lv.g, U, v.R, F, v.AD, lv.mask, p.omega, p.prerelax, 0, ctl, H.nb);
    }
}

// n damped-Jacobi sweeps on level l ending in U; `pre` forms the first
// sweep's input as prerelax*F/AD (preRelc, mgcg1.cu:1190-1197); `in_tmp`:
// the input iterate is in the level's Tmp (an odd n then needs no copy).
template <class T>
void sweeps_t(const gputo_handle *h, const HierView<T> &H, int l, const T *F, T *U, int n, bool pre, const int *ctl,
              bool in_tmp = false)
{
    if (H.per) return sweeps_per(h, H, l, F, U, n, pre, ctl);
    const LevelView<T> &v = H.L[l];
    if (pre && l == 0 && n > 0 && !H.mg9 && H.l0 >= L0_MFMA_F32) {
        // MFMA level 0: the prerelaxed iterate is written by an elementwise
        // pass and the first sweep is a plain one (measured at 192x96x96:
        // 11 + 64 us against 93 us for the fused form, which stages F and AD
        // for every halo node; the FP64 node-gather keeps its fused form,
        // 87 us against 22 + 70 us).  It goes where the sweeps end in U.
        in_tmp = n % 2 == 1;
        if (!H.pre0_ready)
            launch_prerelax<T>(h->stream, F, v.AD, in_tmp ? v.Tmp : U, h->L[l].g.ndof, preset_of(h->prm).prerelax,
                               ctl, H.nb);
        pre = false;
    }
    const T *src = in_tmp ? v.Tmp : U;
    for (int s = 0; s < n; ++s) {
        T *dst = ((n - 1 - s) % 2 == 0) ? U : v.Tmp;
        if (!(pre && s == 0) && dst == src) {
            launch_copy<T, T>(h->stream, U, v.Tmp, h->L[l].g.ndof, ctl, H.nb);
            src = v.Tmp;
        }
        OpArgsT<T> a = view_args(h, H, l, ctl);
        a.U = src;
        a.F = F;
        a.out = dst;
        level_op(h, H, l, (pre && s == 0) ? OP_PRE_JACOBI : OP_JACOBI, a);
        src = dst;
    }
}

// Level-l sweeps of the FP32 V-cycle set by the params (0: the preset's).
inline int mixed_sweeps(const gputo_params &prm, int l, bool post)
{
    if (l == 0) return post ? prm.mixed_npost0 : prm.mixed_npre0;
    if (l == 1) return post ? prm.mixed_npost1 : prm.mixed_npre1;
    if (l >= 3) {
        const int n = post ? prm.mixed_npost3 : prm.mixed_npre3;
        if (n != 0) return n;
    }
    return post ? prm.mixed_npost2 : prm.mixed_npre2;
}

// Sweeps before / after the coarse correction on level l: the preset's,
// or in the FP32 V-cycle the mixed_npre/npost params of level 0, level 1,
// level 2 and the levels from 3 on (-1: none; before the correction the
// prerelaxed iterate alone then smooths the level).
template <class T>
int sweeps_of(const gputo_handle *h, int l, bool post)
{
    const Preset p = preset_of(h->prm);
    if (std::is_same<T, float>::value) {
        const int n = mixed_sweeps(h->prm, l, post);
        if (n > 0) return n;
        if (n == -1) return 0;
    }
    return post ? p.npost : p.npre;
}

// U0 = V(F0).
template <class T>
void vcycle_t(const gputo_handle *h, const HierView<T> &H, const T *F0, T *U0, const int *ctl)
{
    const Preset p = preset_of(h->prm);
    const int nl = h->nl;
    for (int l = 0; l < nl; ++l) {
        const LevelView<T> &v = H.L[l];
        const Level &lv = h->L[l];
        const T *F = l == 0 ? F0 : v.F;
        T *U = l == 0 ? U0 : v.U;
        if (l > 0) {
            launch_restrict<T>(h->stream, lv.g, H.L[l - 1].R, v.F, lv.mask, ctl, H.nb);
            if (H.per) launch_per_fold<T>(h->stream, lv.g, v.F, 0, ctl, H.nb);
        }
        if (l < nl - 1) {
            const int npre = sweeps_of<T>(h, l, false);
            if (npre == 0 && !H.per) {
                OpArgsT<T> a = view_args(h, H, l, ctl);
                a.F = F;
                a.out = v.R;
                a.out2 = U;
                level_op(h, H, l, OP_PRE_DEFECT, a);
            } else {
                if (npre == 0)
                    launch_per_jacobi<T>(h->stream, lv.g, U, v.R, F, v.AD, lv.mask, p.omega, p.prerelax, 1, ctl, H.nb);
                else
                    sweeps_t(h, H, l, F, U, npre, true, ctl);
                OpArgsT<T> a = view_args(h, H, l, ctl);
                a.U = U;
                a.F = F;
                a.out = v.R;
                level_op(h, H, l, OP_DEFECT, a);
                if (H.per) launch_per_fold<T>(h->stream, lv.g, v.R, 0, ctl, H.nb);
            }
        } else if (!h->C.direct) {
            // coarsest level too large to factor: 10*npre sweeps from the
            // level's previous iterate (mgcg9.cu:553-563)
            sweeps_t(h, H, l, F, U, 10 * p.npre, false, ctl);
        } else {
            launch_coarse_solve<T>(h->stream, h->C, F, U, lv.mask, ctl, H.nb);
            if (H.per) launch_per_bcast<T>(h->stream, lv.g, U, ctl, H.nb);
        }
    }
    for (int l = nl - 2; l >= 0; --l) {
        const LevelView<T> &v = H.L[l];
        const T *F = l == 0 ? F0 : v.F;
        T *U = l == 0 ? U0 : v.U;
        // an odd number of post-sweeps starts from Tmp so the last one
        // writes U without a copy
        const int npost = sweeps_of<T>(h, l, true);
        const bool to_tmp = !H.per && npost % 2 == 1;
        launch_interp_to<T>(h->stream, h->L[l].g, H.L[l + 1].U, U, to_tmp ? H.L[l].Tmp : U, h->L[l].mask,
                            p.zero_after_interp, ctl, H.nb);
        sweeps_t(h, H, l, F, U, npost, false, ctl, to_tmp);
        if constexpr (std::is_same<T, float>::value) {
            if (H.learn && l == kLearnLevel) learn_stage(const_cast<gputo_handle *>(h), H, l, F, U, ctl);
        }
    }
}

}  // namespace gputo
