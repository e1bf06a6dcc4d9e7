/*
 * gputo_mex.h -- handle cache shared by the MEX shims mgcg1_gputo.c and
 * mgcg9_gputo.c.  The reference MEX functions set up and tear down their
 * device state on every call (Cantilever/TOP3D125MGCGCUDA/mgcg1.cu:184-233
 * with cudaDeviceReset at :202); the shims keep one libgputo handle per
 * (mesh, levels, preset, sweeps) across calls instead and destroy the
 * handles when MATLAB unloads the MEX file (mexAtExit).
 */
#ifndef GPUTO_MEX_H
#define GPUTO_MEX_H

#include "gputo.h"
#include "mex.h"

#define GPUTO_MEX_CACHE 8

struct gputo_mex_slot {
    int key[6];
    gputo_handle *h;
};
static struct gputo_mex_slot gputo_mex_cache[GPUTO_MEX_CACHE];

static void gputo_mex_release(void)
{
    for (int i = 0; i < GPUTO_MEX_CACHE; ++i)
        if (gputo_mex_cache[i].h) {
            gputo_destroy(gputo_mex_cache[i].h);
            gputo_mex_cache[i].h = 0;
        }
}

/* The cached handle of this mesh and engine (created on first use; the
 * oldest slot is recycled when all are taken); NULL when creation fails. */
static gputo_handle *gputo_mex_handle(int nelx, int nely, int nelz, int nl, int preset, int nswp, int print_level)
{
    const int key[6] = {nelx, nely, nelz, nl, preset, nswp};
    int slot = -1;
    for (int i = 0; i < GPUTO_MEX_CACHE; ++i) {
        if (gputo_mex_cache[i].h) {
            int same = 1;
            for (int k = 0; k < 6; ++k) same = same && gputo_mex_cache[i].key[k] == key[k];
            if (same) return gputo_mex_cache[i].h;
        } else if (slot < 0) {
            slot = i;
        }
    }
    if (slot < 0) {
        gputo_destroy(gputo_mex_cache[0].h);
        gputo_mex_cache[0].h = 0;
        slot = 0;
    }
    gputo_params p;
    gputo_default_params(&p);
    p.preset = preset;
    p.nswp = nswp;
    p.print_level = print_level;
    gputo_handle *h = 0;
    if (gputo_create(nelx, nely, nelz, nl, &p, &h) != 0) return 0;
    for (int k = 0; k < 6; ++k) gputo_mex_cache[slot].key[k] = key[k];
    gputo_mex_cache[slot].h = h;
    mexAtExit(gputo_mex_release);
    return h;
}

#endif
