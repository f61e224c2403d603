// join_dispatch.h -- expands the (KC, R) instantiations of one dtype into a
// launcher, an availability query, an LDS-size query and an occupancy query.
// Included once per dtype translation unit with BFJ_TR (traits), BFJ_SFX (name
// suffix) and BFJ_LIST(X) (the X(KC, R) pairs that are built) defined.
#pragma once
#include "join_kernel.h"

#define BFJ_CAT2(a, b) a##b
#define BFJ_CAT(a, b) BFJ_CAT2(a, b)

int BFJ_CAT(bfj_join_launch_, BFJ_SFX)(const JoinLaunch &L, const JoinArgs &a, hipStream_t s)
{
    const unsigned grid = L.grid_blocks > 0 ? (unsigned)L.grid_blocks : (unsigned)a.n_items;
#define BFJ_X(KC_, R_)                                                                                       \
    if (L.kc == KC_ && L.rows == R_) {                                                                        \
        if (hipFuncSetAttribute((const void *)k_join<BFJ_TR, KC_, R_>,                                        \
                                hipFuncAttributeMaxDynamicSharedMemorySize, L.lds_bytes) != hipSuccess)      \
            return 2;                                                                                         \
        hipLaunchKernelGGL((k_join<BFJ_TR, KC_, R_>), dim3(grid), dim3(BFJ_JBLOCK), L.lds_bytes, s, a);        \
        return hipGetLastError() == hipSuccess ? 0 : 2;                                                       \
    }
    BFJ_LIST(BFJ_X)
#undef BFJ_X
    return 1;
}

bool BFJ_CAT(bfj_join_has_, BFJ_SFX)(int kc, int rows)
{
#define BFJ_X(KC_, R_) \
    if (kc == KC_ && rows == R_) return true;
    BFJ_LIST(BFJ_X)
#undef BFJ_X
    return false;
}

int BFJ_CAT(bfj_join_lds_, BFJ_SFX)(int kc, int rows, int n_eps)
{
#define BFJ_X(KC_, R_) \
    if (kc == KC_ && rows == R_) return JoinLds<BFJ_TR, KC_, R_>::total(n_eps);
    BFJ_LIST(BFJ_X)
#undef BFJ_X
    return 0;
}

int BFJ_CAT(bfj_join_occ_, BFJ_SFX)(int kc, int rows, int lds_bytes)
{
    int nb = 0;
#define BFJ_X(KC_, R_)                                                                                  \
    if (kc == KC_ && rows == R_) {                                                                       \
        if (hipFuncSetAttribute((const void *)k_join<BFJ_TR, KC_, R_>,                                   \
                                hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes) != hipSuccess)   \
            return 0;                                                                                    \
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (k_join<BFJ_TR, KC_, R_>), BFJ_JBLOCK,     \
                                                         lds_bytes) != hipSuccess)                      \
            return 0;                                                                                    \
        return nb;                                                                                       \
    }
    BFJ_LIST(BFJ_X)
#undef BFJ_X
    return 0;
}
