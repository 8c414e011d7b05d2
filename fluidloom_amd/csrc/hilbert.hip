// a1/a2: Hilbert key kernels.  One thread per cell in a grid-stride loop;
// SoA int32 coordinates in, u64 keys out (21 B per cell of HBM traffic for
// the per-element-level form).  The 12-state table is expanded into LDS once
// per workgroup (hilbert_tables.hpp).
#include "fl_internal.hpp"
#include "hilbert_tables.hpp"

namespace fl {
namespace {

constexpr int kBlock = 256;

unsigned stride_grid(size_t n) {
    // enough workgroups for ~8 per CU on 256 CUs, never more than the work
    size_t g = (n + kBlock - 1) / kBlock;
    if (g > 2048) g = 2048;
    return (unsigned)(g ? g : 1);
}

__global__ __launch_bounds__(kBlock) void k_encode(const int32_t* __restrict__ x,
                                                   const int32_t* __restrict__ y,
                                                   const int32_t* __restrict__ z,
                                                   const uint8_t* __restrict__ lev, int lev_scalar,
                                                   int mode, uint64_t* __restrict__ out, size_t n,
                                                   size_t stride) {
    // mode 0: raw encode; 1: key63; 2: key63 | level
    __shared__ HilbertLds L;
    hilbert_lds_init(L, false);
    __syncthreads();
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        int l = lev ? (int)lev[i] : lev_scalar;
        uint64_t k;
        if (l > 21) {
            k = kEmptyKey;
        } else {
            k = hilbert_encode_lds(L, (uint32_t)x[i * stride], (uint32_t)y[i * stride], (uint32_t)z[i * stride], l);
            if (mode == 1) k <<= 3 * (21 - l);
            else if (mode == 2) k = (k << (3 * (21 - l))) | (uint64_t)l;
        }
        out[i] = k;
    }
}

__global__ __launch_bounds__(kBlock) void k_decode(const uint64_t* __restrict__ key,
                                                   const uint8_t* __restrict__ lev, int lev_scalar,
                                                   int32_t* __restrict__ x, int32_t* __restrict__ y,
                                                   int32_t* __restrict__ z, size_t n, size_t stride) {
    __shared__ HilbertLds L;
    hilbert_lds_init(L, true);
    __syncthreads();
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        int l = lev ? (int)lev[i] : lev_scalar;
        int32_t a = 0, b = 0, c = 0;
        if (l > 0 && l <= 21) hilbert_decode_lds(L, key[i], l, &a, &b, &c);
        x[i * stride] = a; y[i * stride] = b; z[i * stride] = c;
    }
}

__global__ __launch_bounds__(kBlock) void k_hkey(const int32_t* __restrict__ x, const int32_t* __restrict__ y,
                                                 const int32_t* __restrict__ z, const uint8_t* __restrict__ lev,
                                                 uint64_t* __restrict__ out, size_t n) {
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const int32_t a = x[i], b = y[i], c = z[i];
        const int l = lev[i];
        out[i] = (l <= 19 && a >= 0 && b >= 0 && c >= 0 && a < kKeyCoordLimit && b < kKeyCoordLimit &&
                  c < kKeyCoordLimit)
                     ? cell_hkey((uint32_t)a, (uint32_t)b, (uint32_t)c, l)
                     : kEmptyKey;
    }
}

// adaptation::hilbert_encode_3d: bits max_level-1..0 of the raw coordinates,
// no masking, rotation/direction tables.
__global__ __launch_bounds__(kBlock) void k_rotdir(const int32_t* __restrict__ x,
                                                   const int32_t* __restrict__ y,
                                                   const int32_t* __restrict__ z, int max_level,
                                                   uint64_t* __restrict__ out, size_t n) {
    __shared__ uint8_t rot[64], dir[64];
    if (threadIdx.x < 64) {
        rot[threadIdx.x] = kROT[threadIdx.x >> 3][threadIdx.x & 7];
        dir[threadIdx.x] = kDIR[threadIdx.x >> 3][threadIdx.x & 7];
    }
    __syncthreads();
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        int32_t a = x[i], b = y[i], c = z[i];
        uint64_t h = 0;
        int r = 0;
        for (int l = max_level - 1; l >= 0; --l) {
            int blk = ((a >> l) & 1) | (((b >> l) & 1) << 1) | (((c >> l) & 1) << 2);
            h = (h << 3) | rot[r * 8 + blk];
            r = dir[r * 8 + blk];
        }
        out[i] = h;
    }
}

}  // namespace
}  // namespace fl

using namespace fl;

extern "C" {

fl_status fl_hilbert_encode(const int32_t* d_x, const int32_t* d_y, const int32_t* d_z,
                            const uint8_t* d_level, int level_scalar, uint64_t* d_key, size_t n,
                            fl_stream s) {
    if (n == 0) return FL_OK;
    FL_REQUIRE(d_x && d_y && d_z && d_key, "null pointer");
    FL_REQUIRE(d_level || (level_scalar >= 0 && level_scalar <= 21), "level must be in 0..21");
    hipLaunchKernelGGL(k_encode, dim3(stride_grid(n)), dim3(kBlock), 0, FL_STREAM(s), d_x, d_y, d_z,
                       d_level, level_scalar, 0, d_key, n, (size_t)1);
    FL_CHECK_LAUNCH();
    return FL_OK;
}

fl_status fl_hilbert_decode(const uint64_t* d_key, const uint8_t* d_level, int level_scalar,
                            int32_t* d_x, int32_t* d_y, int32_t* d_z, size_t n, fl_stream s) {
    if (n == 0) return FL_OK;
    FL_REQUIRE(d_x && d_y && d_z && d_key, "null pointer");
    FL_REQUIRE(d_level || (level_scalar >= 0 && level_scalar <= 21), "level must be in 0..21");
    hipLaunchKernelGGL(k_decode, dim3(stride_grid(n)), dim3(kBlock), 0, FL_STREAM(s), d_key, d_level,
                       level_scalar, d_x, d_y, d_z, n, (size_t)1);
    FL_CHECK_LAUNCH();
    return FL_OK;
}

fl_status fl_hilbert_key63(const int32_t* d_x, const int32_t* d_y, const int32_t* d_z,
                           const uint8_t* d_level, int pack_level, uint64_t* d_key, size_t n,
                           fl_stream s) {
    if (n == 0) return FL_OK;
    FL_REQUIRE(d_x && d_y && d_z && d_key && d_level, "null pointer");
    hipLaunchKernelGGL(k_encode, dim3(stride_grid(n)), dim3(kBlock), 0, FL_STREAM(s), d_x, d_y, d_z,
                       d_level, 0, pack_level ? 2 : 1, d_key, n, (size_t)1);
    FL_CHECK_LAUNCH();
    return FL_OK;
}

fl_status fl_hilbert_encode_aos4(const int32_t* d_xyzw, int level, uint64_t* d_key, size_t n, fl_stream s) {
    if (n == 0) return FL_OK;
    FL_REQUIRE(d_xyzw && d_key, "null pointer");
    FL_REQUIRE(level >= 0 && level <= 21, "level must be in 0..21");
    hipLaunchKernelGGL(k_encode, dim3(stride_grid(n)), dim3(kBlock), 0, FL_STREAM(s), d_xyzw, d_xyzw + 1,
                       d_xyzw + 2, (const uint8_t*)nullptr, level, 0, d_key, n, (size_t)4);
    FL_CHECK_LAUNCH();
    return FL_OK;
}

fl_status fl_hilbert_decode_aos4(const uint64_t* d_key, int level, int32_t* d_xyzw, size_t n, fl_stream s) {
    if (n == 0) return FL_OK;
    FL_REQUIRE(d_xyzw && d_key, "null pointer");
    FL_REQUIRE(level >= 0 && level <= 21, "level must be in 0..21");
    hipLaunchKernelGGL(k_decode, dim3(stride_grid(n)), dim3(kBlock), 0, FL_STREAM(s), d_key,
                       (const uint8_t*)nullptr, level, d_xyzw, d_xyzw + 1, d_xyzw + 2, n, (size_t)4);
    FL_CHECK_LAUNCH();
    return FL_OK;
}

fl_status fl_cell_hkey(const int32_t* d_x, const int32_t* d_y, const int32_t* d_z,
                       const uint8_t* d_level, uint64_t* d_key, size_t n, fl_stream s) {
    if (n == 0) return FL_OK;
    FL_REQUIRE(d_x && d_y && d_z && d_key && d_level, "null pointer");
    hipLaunchKernelGGL(k_hkey, dim3(stride_grid(n)), dim3(kBlock), 0, FL_STREAM(s), d_x, d_y, d_z, d_level,
                       d_key, n);
    FL_CHECK_LAUNCH();
    return FL_OK;
}

fl_status fl_rotdir_encode(const int32_t* d_x, const int32_t* d_y, const int32_t* d_z, int max_level,
                           uint64_t* d_key, size_t n, fl_stream s) {
    if (n == 0) return FL_OK;
    FL_REQUIRE(d_x && d_y && d_z && d_key, "null pointer");
    FL_REQUIRE(max_level >= 0 && max_level <= 21, "max_level must be in 0..21");
    hipLaunchKernelGGL(k_rotdir, dim3(stride_grid(n)), dim3(kBlock), 0, FL_STREAM(s), d_x, d_y, d_z,
                       max_level, d_key, n);
    FL_CHECK_LAUNCH();
    return FL_OK;
}

}  // extern "C"
