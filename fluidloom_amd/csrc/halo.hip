// a14: halo pack/unpack and generic coarse<->fine field transfer.
//
// pack_fields/unpack_fields (kernels/halo_pack.cl:3-45) copy
// packed[offset + 4k] = field[k] (identity index); here an optional index list
// turns them into the gather/scatter a one-GPU ghost exchange needs.  The
// coarse->fine transfer copies the parent value to each child
// (interpolate_split_fields, split_cells.cl:110-132); the fine->coarse
// transfer sums the 8 children in sibling order, scaled by 1/8 for the
// arithmetic rule (merge_fields, merge_cells.cl:123-159).
#include "fl_internal.hpp"

namespace fl {
namespace {

constexpr int kBlock = 256;

unsigned hgrid(size_t n) {
    size_t g = (n + kBlock - 1) / kBlock;
    if (g > 8192) g = 8192;
    return (unsigned)(g ? g : 1);
}

__global__ void k_pack(const float* __restrict__ field, const uint32_t* __restrict__ idx, size_t n,
                       float* __restrict__ out) {
    for (size_t k = (size_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += (size_t)gridDim.x * kBlock)
        out[k] = field[idx ? idx[k] : k];
}

__global__ void k_unpack(const float* __restrict__ in, const uint32_t* __restrict__ idx, size_t n,
                         float* __restrict__ field) {
    for (size_t k = (size_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += (size_t)gridDim.x * kBlock)
        field[idx ? idx[k] : k] = in[k];
}

__global__ void k_c2f(const float* __restrict__ coarse, const uint32_t* __restrict__ parent, size_t nfine,
                      int ncomp, float* __restrict__ fine) {
    const size_t total = nfine * (size_t)ncomp;
    for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < total; e += (size_t)gridDim.x * kBlock) {
        size_t k = e / ncomp;
        int c = (int)(e % ncomp);
        fine[e] = coarse[(size_t)parent[k] * ncomp + c];
    }
}

__global__ void k_f2c(const float* __restrict__ fine, const uint32_t* __restrict__ child8, size_t ncoarse,
                      int ncomp, int rule, float* __restrict__ coarse) {
    const size_t total = ncoarse * (size_t)ncomp;
    for (size_t e = (size_t)blockIdx.x * kBlock + threadIdx.x; e < total; e += (size_t)gridDim.x * kBlock) {
        size_t g = e / ncomp;
        int c = (int)(e % ncomp);
        float s = 0.0f;
        for (int q = 0; q < 8; ++q) {
            float v = fine[(size_t)child8[8 * g + q] * ncomp + c];
            s += rule == 0 ? v * 0.125f : v;
        }
        coarse[e] = s;
    }
}

}  // namespace
}  // namespace fl

using namespace fl;

extern "C" {

fl_status fl_pack_fields(const float* d_field, const uint32_t* d_idx, size_t count, uint8_t* d_packed,
                         size_t offset_bytes, fl_stream s) {
    if (count == 0) return FL_OK;
    FL_REQUIRE(d_field && d_packed, "null pointer");
    FL_REQUIRE(offset_bytes % 4 == 0, "offset must be a multiple of 4 bytes");
    hipLaunchKernelGGL(k_pack, dim3(hgrid(count)), dim3(kBlock), 0, FL_STREAM(s), d_field, d_idx, count,
                       reinterpret_cast<float*>(d_packed + offset_bytes));
    FL_CHECK_LAUNCH();
    return FL_OK;
}

fl_status fl_unpack_fields(const uint8_t* d_packed, size_t offset_bytes, const uint32_t* d_idx, size_t count,
                           float* d_field, fl_stream s) {
    if (count == 0) return FL_OK;
    FL_REQUIRE(d_field && d_packed, "null pointer");
    FL_REQUIRE(offset_bytes % 4 == 0, "offset must be a multiple of 4 bytes");
    hipLaunchKernelGGL(k_unpack, dim3(hgrid(count)), dim3(kBlock), 0, FL_STREAM(s),
                       reinterpret_cast<const float*>(d_packed + offset_bytes), d_idx, count, d_field);
    FL_CHECK_LAUNCH();
    return FL_OK;
}

fl_status fl_ghost_fill_c2f(const float* d_coarse, const uint32_t* d_parent, size_t nfine, int ncomp,
                            float* d_fine, fl_stream s) {
    if (nfine == 0 || ncomp == 0) return FL_OK;
    FL_REQUIRE(d_coarse && d_parent && d_fine && ncomp > 0, "bad argument");
    hipLaunchKernelGGL(k_c2f, dim3(hgrid(nfine * ncomp)), dim3(kBlock), 0, FL_STREAM(s), d_coarse, d_parent, nfine,
                       ncomp, d_fine);
    FL_CHECK_LAUNCH();
    return FL_OK;
}

fl_status fl_ghost_fill_f2c(const float* d_fine, const uint32_t* d_child8, size_t ncoarse, int ncomp, int rule,
                            float* d_coarse, fl_stream s) {
    if (ncoarse == 0 || ncomp == 0) return FL_OK;
    FL_REQUIRE(d_fine && d_child8 && d_coarse && ncomp > 0, "bad argument");
    FL_REQUIRE(rule == 0 || rule == 1, "rule must be 0 (mean) or 1 (sum)");
    hipLaunchKernelGGL(k_f2c, dim3(hgrid(ncoarse * ncomp)), dim3(kBlock), 0, FL_STREAM(s), d_fine, d_child8,
                       ncoarse, ncomp, rule, d_coarse);
    FL_CHECK_LAUNCH();
    return FL_OK;
}

}  // extern "C"
