// mfma_ops.h -- the gfx950 matrix instructions of the join, one traits struct per
// dot-product dtype: operand fragment type, accumulator type, the MFMA itself and
// where its output rows land.  Shared by the join (join_kernel.h) and the accuracy
// statistics (errstat.hip), which must replay the join's arithmetic exactly.
//
// A packed fragment (k_quantize / k_quantize16) is one K-chunk of 16 rows: 64 lane
// operands in lane order.  The join's LDS ring holds fragments in 16-byte halves
// instead: the first 16 bytes of every lane's operand in lane order, then (FP8, whose
// operand is 32 bytes) the second 16 bytes of every lane's, so that each 16-byte LDS
// read of a fragment is lane-contiguous and meets no bank conflict (a 32-byte lane
// stride conflicts 2-way).  load() reads a lane's operand from a packed fragment,
// load_ring() from a ring fragment; both take the fragment's first byte.
#pragma once
#include "bfj_internal.h"

template <class F>
__device__ inline F bfj_frag_load16(const unsigned char *p, int lane)
{
    return reinterpret_cast<const F *>(p)[lane];
}

struct TrF16 {
    static constexpr int dtype = BFJ_FP16;
    typedef bfj_h8 frag_t;
    typedef bfj_f4 acc_t;
    typedef float val_t;
    __device__ static inline frag_t load(const unsigned char *p, int lane) { return bfj_frag_load16<frag_t>(p, lane); }
    __device__ static inline frag_t load_ring(const unsigned char *p, int lane) { return load(p, lane); }
    __device__ static inline acc_t mma(frag_t a, frag_t b, acc_t c, int, int)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
    // Output register r of lane l holds tile row 4*(l/16)+r, column l%16.
    __device__ static inline int out_row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

struct TrBF16 {
    static constexpr int dtype = BFJ_BF16;
    typedef bfj_b8 frag_t;
    typedef bfj_f4 acc_t;
    typedef float val_t;
    __device__ static inline frag_t load(const unsigned char *p, int lane) { return bfj_frag_load16<frag_t>(p, lane); }
    __device__ static inline frag_t load_ring(const unsigned char *p, int lane) { return load(p, lane); }
    __device__ static inline acc_t mma(frag_t a, frag_t b, acc_t c, int, int)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
    __device__ static inline int out_row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

struct TrF8 {
    static constexpr int dtype = BFJ_FP8;
    typedef bfj_i8 frag_t;
    typedef bfj_f4 acc_t;
    typedef float val_t;
    __device__ static inline frag_t load(const unsigned char *p, int lane) { return bfj_frag_load16<frag_t>(p, lane); }
    // Bytes 0-15 of the lane's 32 from the fragment's first half, 16-31 from its second.
    __device__ static inline frag_t load_ring(const unsigned char *p, int lane)
    {
        const bfj_i4 lo = bfj_frag_load16<bfj_i4>(p, lane), hi = bfj_frag_load16<bfj_i4>(p + 16 * BFJ_WAVE, lane);
        return frag_t{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
    }
    // Operand format 0 = FP8 e4m3 for A and B; sa/sb are E8M0 row scales.
    __device__ static inline acc_t mma(frag_t a, frag_t b, acc_t c, int sa, int sb)
    {
        return __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a, b, c, 0, 0, 0, sa, 0, sb);
    }
    __device__ static inline int out_row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

struct TrF64 {
    static constexpr int dtype = BFJ_FP64;
    typedef bfj_d2 frag_t; // two K=4 steps of v_mfma_f64_16x16x4
    typedef bfj_d4 acc_t;
    typedef double val_t;
    __device__ static inline frag_t load(const unsigned char *p, int lane) { return bfj_frag_load16<frag_t>(p, lane); }
    __device__ static inline frag_t load_ring(const unsigned char *p, int lane) { return load(p, lane); }
    __device__ static inline acc_t mma(frag_t a, frag_t b, acc_t c, int, int)
    {
        c = __builtin_amdgcn_mfma_f64_16x16x4f64(a[0], b[0], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f64_16x16x4f64(a[1], b[1], c, 0, 0, 0);
        return c;
    }
    // The FP64 shape stores tile row 4*r + l/16 in register r of lane l.
    __device__ static inline int out_row(int lane, int r) { return 4 * r + (lane >> 4); }
};

