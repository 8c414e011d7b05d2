// a14/a15: D3Q19 lattice kernels for one refinement level.
//
//   k_collide   fused pull-stream + BGK over the active bricks: one 64-lane
//               wave per 4x4x4 brick, lane = cell.  The brick's 27-entry
//               neighbour row sits in lanes 0..26 and is fetched with
//               __shfl, so a regular pull costs one HBM-streamed population
//               load; cells of "simple" bricks (all 64 active, no wall links)
//               never read their link word.  Workgroups are remapped so that
//               each XCD processes one contiguous Hilbert-ordered range of
//               bricks, keeping cross-brick re-reads inside one L2.
//   k_capture   pre-collision state of coarse cells that feed c2f ghosts
//   k_fill_c2f  coarse -> fine ghost populations (f_neq rescaled)
//   k_fill_f2c  fine -> coarse ghost populations (8-child average)
//   k_init, k_totals, k_force, k_extract: set-up, FP64 reductions, queries
#include "fl_internal.hpp"
#include "lbm_device.hpp"
#include "lbm_kernels.hpp"
#include "wave.hpp"

namespace fl {
namespace lbm {
namespace {

constexpr int kWG = 256;
constexpr int kBricksPerWG = kWG / 64;

__device__ inline unsigned xcd_remap(unsigned bid, unsigned nblk) {
    const unsigned q = nblk / 8, r = nblk % 8;
    const unsigned x = bid % 8, k = bid / 8;
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

template <typename T>
__global__ __launch_bounds__(kWG) void k_collide(LevelDev L, BCDev bc, const T* __restrict__ A, T* __restrict__ B,
                                                 typename Store<T>::R inv_tau) {
    using R = typename Store<T>::R;
    const unsigned vb = xcd_remap(blockIdx.x, gridDim.x);
    const int b = (int)(vb * kBricksPerWG + (threadIdx.x >> 6));
    if (b >= L.nactive) return;  // wave-uniform
    const int lane = (int)(threadIdx.x & 63);
    const int nbrv = lane < 27 ? L.nbr[b * 27 + lane] : -1;
    const bool simple = (L.bflags[b] & 1u) != 0;
    const size_t slot = (size_t)b * 64 + lane;
    const uint64_t w = simple ? 0ull : L.cflags[slot];
    const bool active = simple || (w & kActiveBit);
    const int lx = lane & 3, ly = (lane >> 2) & 3, lz = lane >> 4;
    const size_t S = L.S;
    R f[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const int px = lx - cx(i), py = ly - cy(i), pz = lz - cz(i);
        const int nbi = 13 + (px >> 2) + 3 * (py >> 2) + 9 * (pz >> 2);
        const int nb = __shfl(nbrv, nbi, 64);
        const size_t src = (size_t)nb * 64 + (size_t)((px & 3) + 4 * (py & 3) + 16 * (pz & 3));
        f[i] = 0;
        if (active) {
            const unsigned code = (unsigned)(w >> (2 * i)) & 3u;
            if (code == kLinkRegular) {
                f[i] = Store<T>::load(A, (size_t)i * S + src, i);
            } else {
                if (code == kLinkSolid) {
                    f[i] = Store<T>::load(A, (size_t)opp(i) * S + slot, opp(i));
                } else if (code == kLinkWall) {
                    const int32_t* c3 = L.bcoord + 3 * b;
                    int face = wall_face(bc, L.ext, c3[0] * 4 + px, c3[1] * 4 + py, c3[2] * 4 + pz);
                    R uw = 0;
                    if (face >= 0) {
                        if (sizeof(R) == 8)
                            uw = (R)(cx(i) * bc.ud[face][0] + cy(i) * bc.ud[face][1] + cz(i) * bc.ud[face][2]);
                        else
                            uw = (R)((float)cx(i) * bc.u[face][0] + (float)cy(i) * bc.u[face][1] +
                                     (float)cz(i) * bc.u[face][2]);
                    }
                    f[i] = Store<T>::load(A, (size_t)opp(i) * S + slot, opp(i)) + (R)6 * (R)wq(i) * (R)bc.rho0 * uw;
                } else {
                    const int32_t* c3 = L.bcoord + 3 * b;
                    int gx = c3[0] * 4 + px, gy = c3[1] * 4 + py, gz = c3[2] * 4 + pz;
                    int qx = gx < 0 ? 0 : (gx >= L.ext[0] ? L.ext[0] - 1 : gx);
                    int qy = gy < 0 ? 0 : (gy >= L.ext[1] ? L.ext[1] - 1 : gy);
                    int qz = gz < 0 ? 0 : (gz >= L.ext[2] ? L.ext[2] - 1 : gz);
                    int64_t s2 = slot_of(L, b, qx - c3[0] * 4, qy - c3[1] * 4, qz - c3[2] * 4);
                    f[i] = Store<T>::load(A, (size_t)i * S + (size_t)s2, i);
                }
            }
        }
    }
    if (!active) return;
    bgk(f, inv_tau);
#pragma unroll
    for (int i = 0; i < Q; ++i) Store<T>::store(B, (size_t)i * S + slot, i, f[i]);
}

template <typename T>
__global__ __launch_bounds__(kWG) void k_capture(LevelDev L, BCDev bc, const T* __restrict__ A,
                                                 const uint32_t* __restrict__ cap_slots, uint32_t ncap,
                                                 typename Store<T>::R* __restrict__ cap) {
    using R = typename Store<T>::R;
    for (uint32_t k = blockIdx.x * kWG + threadIdx.x; k < ncap; k += gridDim.x * kWG) {
        R f[Q];
        gather_slot<T, R>(L, bc, A, cap_slots[k], f);
#pragma unroll
        for (int i = 0; i < Q; ++i) cap[(size_t)i * ncap + k] = f[i];
    }
}

template <typename T>
__global__ __launch_bounds__(kWG) void k_fill_c2f(const uint32_t* __restrict__ ghost, const uint32_t* __restrict__ capidx,
                                                  uint32_t n, const typename Store<T>::R* __restrict__ cap0,
                                                  const typename Store<T>::R* __restrict__ cap1, uint32_t ncap,
                                                  typename Store<T>::R alpha, typename Store<T>::R fac,
                                                  T* __restrict__ A, size_t S) {
    using R = typename Store<T>::R;
    for (uint32_t k = blockIdx.x * kWG + threadIdx.x; k < n; k += gridDim.x * kWG) {
        const uint32_t c = capidx[k];
        R f[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i)
            f[i] = ((R)1 - alpha) * cap0[(size_t)i * ncap + c] + alpha * cap1[(size_t)i * ncap + c];
        R rho, ux, uy, uz;
        moments(f, rho, ux, uy, uz);
        const R usq = ux * ux + uy * uy + uz * uz;
        const size_t g = ghost[k];
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const R e = feq(i, rho, ux, uy, uz, usq);
            Store<T>::store(A, (size_t)i * S + g, i, e + fac * (f[i] - e));
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kWG) void k_fill_f2c(const uint32_t* __restrict__ ghost, const uint32_t* __restrict__ child8,
                                                  uint32_t n, LevelDev Lf, BCDev bc, const T* __restrict__ Af,
                                                  typename Store<T>::R fac, T* __restrict__ Ac, size_t Sc) {
    using R = typename Store<T>::R;
    for (uint32_t k = blockIdx.x * kWG + threadIdx.x; k < n; k += gridDim.x * kWG) {
        R sr = 0, sjx = 0, sjy = 0, sjz = 0, sneq[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i) sneq[i] = 0;
        for (int q = 0; q < 8; ++q) {
            R f[Q];
            gather_slot<T, R>(Lf, bc, Af, child8[(size_t)8 * k + q], f);
            R rho, ux, uy, uz;
            moments(f, rho, ux, uy, uz);
            const R usq = ux * ux + uy * uy + uz * uz;
            sr += rho;
            sjx += rho * ux;
            sjy += rho * uy;
            sjz += rho * uz;
#pragma unroll
            for (int i = 0; i < Q; ++i) sneq[i] += f[i] - feq(i, rho, ux, uy, uz, usq);
        }
        const R r = sr * (R)0.125;
        const R ux = (sjx * (R)0.125) / r, uy = (sjy * (R)0.125) / r, uz = (sjz * (R)0.125) / r;
        const R usq = ux * ux + uy * uy + uz * uz;
        const size_t g = ghost[k];
#pragma unroll
        for (int i = 0; i < Q; ++i)
            Store<T>::store(Ac, (size_t)i * Sc + g, i, feq(i, r, ux, uy, uz, usq) + fac * (sneq[i] * (R)0.125));
    }
}

template <typename T>
__global__ __launch_bounds__(kWG) void k_init(size_t S, double rho, double ux, double uy, double uz,
                                              T* __restrict__ A, T* __restrict__ B) {
    using R = typename Store<T>::R;
    const R usq = (R)(ux * ux + uy * uy + uz * uz);
    for (size_t s = (size_t)blockIdx.x * kWG + threadIdx.x; s < S; s += (size_t)gridDim.x * kWG) {
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            R e = feq(i, (R)rho, (R)ux, (R)uy, (R)uz, usq);
            Store<T>::store(A, (size_t)i * S + s, i, e);
            Store<T>::store(B, (size_t)i * S + s, i, e);
        }
    }
}

// Per-block partial sums of (mass, jx, jy, jz) over active slots, in FP64.
template <typename T>
__global__ __launch_bounds__(kWG) void k_totals(LevelDev L, const T* __restrict__ A, double vol,
                                                double* __restrict__ partial) {
    using R = typename Store<T>::R;
    double acc[4] = {0, 0, 0, 0};
    const size_t n = (size_t)L.nactive * 64;
    for (size_t s = (size_t)blockIdx.x * kWG + threadIdx.x; s < n; s += (size_t)gridDim.x * kWG) {
        const bool simple = (L.bflags[s >> 6] & 1u) != 0;
        if (!simple && !(L.cflags[s] & kActiveBit)) continue;
        double r = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            double f = (double)Store<T>::load(A, (size_t)i * L.S + s, i);
            r += f;
            jx += cx(i) * f;
            jy += cy(i) * f;
            jz += cz(i) * f;
        }
        acc[0] += r * vol;
        acc[1] += jx * vol;
        acc[2] += jy * vol;
        acc[3] += jz * vol;
    }
    __shared__ double red[4][kWG];
    for (int c = 0; c < 4; ++c) red[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int st = kWG / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int c = 0; c < 4; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 4; ++c) partial[(size_t)blockIdx.x * 4 + c] = red[c][0];
}

// Momentum exchange over solid links (slot * 32 + q): F += -2 c_q f*_opp(q).
template <typename T>
__global__ __launch_bounds__(kWG) void k_force(const uint32_t* __restrict__ links, uint32_t n, size_t S,
                                               const T* __restrict__ A, double scale,
                                               double* __restrict__ partial) {
    double acc[3] = {0, 0, 0};
    for (uint32_t k = blockIdx.x * kWG + threadIdx.x; k < n; k += gridDim.x * kWG) {
        const uint32_t e = links[k];
        const size_t s = e >> 5;
        const int q = (int)(e & 31u);
        double fo = 0;
        // q-dependent opposite load
#pragma unroll
        for (int i = 1; i < Q; ++i)
            if (i == q) fo = (double)Store<T>::load(A, (size_t)opp(i) * S + s, opp(i));
        int c0 = 0, c1 = 0, c2 = 0;
#pragma unroll
        for (int i = 1; i < Q; ++i)
            if (i == q) { c0 = cx(i); c1 = cy(i); c2 = cz(i); }
        acc[0] += -2.0 * c0 * fo * scale;
        acc[1] += -2.0 * c1 * fo * scale;
        acc[2] += -2.0 * c2 * fo * scale;
    }
    __shared__ double red[3][kWG];
    for (int c = 0; c < 3; ++c) red[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int st = kWG / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int c = 0; c < 3; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; ++c) partial[(size_t)blockIdx.x * 3 + c] = red[c][0];
}

// Populations of listed slots as FP64 [k][q].
template <typename T>
__global__ __launch_bounds__(kWG) void k_extract(const T* __restrict__ A, size_t S, const int64_t* __restrict__ slots,
                                                 size_t n, double* __restrict__ out) {
    for (size_t k = (size_t)blockIdx.x * kWG + threadIdx.x; k < n; k += (size_t)gridDim.x * kWG) {
        const int64_t s = slots[k];
#pragma unroll
        for (int i = 0; i < Q; ++i) out[k * Q + i] = s < 0 ? 0.0 : (double)Store<T>::load(A, (size_t)i * S + s, i);
    }
}

template <typename T>
__global__ __launch_bounds__(kWG) void k_inject(T* __restrict__ A, T* __restrict__ B, size_t S,
                                                const int64_t* __restrict__ slots, size_t n,
                                                const double* __restrict__ in) {
    using R = typename Store<T>::R;
    for (size_t k = (size_t)blockIdx.x * kWG + threadIdx.x; k < n; k += (size_t)gridDim.x * kWG) {
        const int64_t s = slots[k];
        if (s < 0) continue;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            Store<T>::store(A, (size_t)i * S + s, i, (R)in[k * Q + i]);
            Store<T>::store(B, (size_t)i * S + s, i, (R)in[k * Q + i]);
        }
    }
}

// |vorticity| (level-0 units) at listed fluid-leaf slots: central
// differences over same-level fluid-leaf neighbours, one-sided at the edge
// of the level, zero when neither side is a fluid leaf.
template <typename T>
__device__ inline bool leaf_velocity(const LevelDev& L, const T* __restrict__ A, int64_t s, double* u) {
    if (s < 0) return false;
    const bool simple = (L.bflags[s >> 6] & 1u) != 0;
    if (!simple && !(L.cflags[s] & kActiveBit)) return false;
    double r = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        double f = (double)Store<T>::load(A, (size_t)i * L.S + (size_t)s, i);
        r += f;
        jx += cx(i) * f;
        jy += cy(i) * f;
        jz += cz(i) * f;
    }
    u[0] = jx / r;
    u[1] = jy / r;
    u[2] = jz / r;
    return true;
}

template <typename T>
__global__ __launch_bounds__(kWG) void k_vorticity(LevelDev L, const T* __restrict__ A, const int64_t* __restrict__ slots,
                                                   size_t n, double scale, double* __restrict__ out) {
    for (size_t k = (size_t)blockIdx.x * kWG + threadIdx.x; k < n; k += (size_t)gridDim.x * kWG) {
        const int64_t s = slots[k];
        double u0[3];
        if (!leaf_velocity<T>(L, A, s, u0)) {
            out[k] = 0.0;
            continue;
        }
        const int b = (int)(s >> 6), l = (int)(s & 63);
        const int lx = l & 3, ly = (l >> 2) & 3, lz = l >> 4;
        double g[3][3];
        for (int a = 0; a < 3; ++a) {
            const int dx = a == 0, dy = a == 1, dz = a == 2;
            double up[3], um[3];
            const bool hp = leaf_velocity<T>(L, A, slot_of(L, b, lx + dx, ly + dy, lz + dz), up);
            const bool hm = leaf_velocity<T>(L, A, slot_of(L, b, lx - dx, ly - dy, lz - dz), um);
            for (int c = 0; c < 3; ++c)
                g[a][c] = hp && hm ? 0.5 * (up[c] - um[c]) : (hp ? up[c] - u0[c] : (hm ? u0[c] - um[c] : 0.0));
        }
        const double wx = g[1][2] - g[2][1], wy = g[2][0] - g[0][2], wz = g[0][1] - g[1][0];
        out[k] = sqrt(wx * wx + wy * wy + wz * wz) * scale;
    }
}

// Population transfer of a regrid: copy / split child / merged parent / rest.
template <typename T>
__global__ __launch_bounds__(kWG) void k_remap_copy(const T* __restrict__ Ao, size_t So, const uint32_t* __restrict__ dst,
                                                    const uint32_t* __restrict__ src, uint32_t n, T* __restrict__ A,
                                                    T* __restrict__ B, size_t S) {
    for (uint32_t k = blockIdx.x * kWG + threadIdx.x; k < n; k += gridDim.x * kWG) {
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const T v = Ao[(size_t)i * So + src[k]];
            A[(size_t)i * S + dst[k]] = v;
            B[(size_t)i * S + dst[k]] = v;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kWG) void k_remap_split(LevelDev Lc, BCDev bc, const T* __restrict__ Acprev,
                                                     const uint32_t* __restrict__ dst, const uint32_t* __restrict__ src,
                                                     uint32_t n, typename Store<T>::R fac, T* __restrict__ A,
                                                     T* __restrict__ B, size_t S) {
    using R = typename Store<T>::R;
    for (uint32_t k = blockIdx.x * kWG + threadIdx.x; k < n; k += gridDim.x * kWG) {
        R f[Q];
        gather_slot<T, R>(Lc, bc, Acprev, src[k], f);
        R rho, ux, uy, uz;
        moments(f, rho, ux, uy, uz);
        const R usq = ux * ux + uy * uy + uz * uz;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const R e = feq(i, rho, ux, uy, uz, usq);
            const R v = e + fac * (f[i] - e);
            Store<T>::store(A, (size_t)i * S + dst[k], i, v);
            Store<T>::store(B, (size_t)i * S + dst[k], i, v);
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kWG) void k_remap_merge(LevelDev Lf, BCDev bc, const T* __restrict__ Afprev,
                                                     const uint32_t* __restrict__ dst, const uint32_t* __restrict__ src8,
                                                     uint32_t n, typename Store<T>::R fac, T* __restrict__ A,
                                                     T* __restrict__ B, size_t S) {
    using R = typename Store<T>::R;
    for (uint32_t k = blockIdx.x * kWG + threadIdx.x; k < n; k += gridDim.x * kWG) {
        R sr = 0, sjx = 0, sjy = 0, sjz = 0, sneq[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i) sneq[i] = 0;
        for (int q = 0; q < 8; ++q) {
            R f[Q];
            gather_slot<T, R>(Lf, bc, Afprev, src8[(size_t)8 * k + q], f);
            R rho, ux, uy, uz;
            moments(f, rho, ux, uy, uz);
            const R usq = ux * ux + uy * uy + uz * uz;
            sr += rho;
            sjx += rho * ux;
            sjy += rho * uy;
            sjz += rho * uz;
#pragma unroll
            for (int i = 0; i < Q; ++i) sneq[i] += f[i] - feq(i, rho, ux, uy, uz, usq);
        }
        const R r = sr * (R)0.125;
        const R ux = (sjx * (R)0.125) / r, uy = (sjy * (R)0.125) / r, uz = (sjz * (R)0.125) / r;
        const R usq = ux * ux + uy * uy + uz * uz;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const R v = feq(i, r, ux, uy, uz, usq) + fac * (sneq[i] * (R)0.125);
            Store<T>::store(A, (size_t)i * S + dst[k], i, v);
            Store<T>::store(B, (size_t)i * S + dst[k], i, v);
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kWG) void k_remap_rest(const uint32_t* __restrict__ dst, uint32_t n, T* __restrict__ A,
                                                    T* __restrict__ B, size_t S) {
    using R = typename Store<T>::R;
    for (uint32_t k = blockIdx.x * kWG + threadIdx.x; k < n; k += gridDim.x * kWG) {
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const R v = (R)wq(i);
            Store<T>::store(A, (size_t)i * S + dst[k], i, v);
            Store<T>::store(B, (size_t)i * S + dst[k], i, v);
        }
    }
}

// Ghost slots after a regrid: equilibrium of the covering coarse leaf (c2f)
// or of the 8 children's average density and momentum (f2c), both buffers.
template <typename T>
__global__ __launch_bounds__(kWG) void k_ghost_eq_c2f(const uint32_t* __restrict__ ghost, const uint32_t* __restrict__ capidx,
                                                      const uint32_t* __restrict__ cap_slots, uint32_t n,
                                                      const T* __restrict__ Ac, size_t Sc, T* __restrict__ A,
                                                      T* __restrict__ B, size_t S) {
    using R = typename Store<T>::R;
    for (uint32_t k = blockIdx.x * kWG + threadIdx.x; k < n; k += gridDim.x * kWG) {
        const size_t cs = cap_slots[capidx[k]];
        R f[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i) f[i] = Store<T>::load(Ac, (size_t)i * Sc + cs, i);
        R rho, ux, uy, uz;
        moments(f, rho, ux, uy, uz);
        const R usq = ux * ux + uy * uy + uz * uz;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const R e = feq(i, rho, ux, uy, uz, usq);
            Store<T>::store(A, (size_t)i * S + ghost[k], i, e);
            Store<T>::store(B, (size_t)i * S + ghost[k], i, e);
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kWG) void k_ghost_eq_f2c(const uint32_t* __restrict__ ghost, const uint32_t* __restrict__ child8,
                                                      uint32_t n, const T* __restrict__ Af, size_t Sf, T* __restrict__ A,
                                                      T* __restrict__ B, size_t S) {
    using R = typename Store<T>::R;
    for (uint32_t k = blockIdx.x * kWG + threadIdx.x; k < n; k += gridDim.x * kWG) {
        R sr = 0, sjx = 0, sjy = 0, sjz = 0;
        for (int q = 0; q < 8; ++q) {
            const size_t cs = child8[(size_t)8 * k + q];
            R f[Q];
#pragma unroll
            for (int i = 0; i < Q; ++i) f[i] = Store<T>::load(Af, (size_t)i * Sf + cs, i);
            R rho, ux, uy, uz;
            moments(f, rho, ux, uy, uz);
            sr += rho;
            sjx += rho * ux;
            sjy += rho * uy;
            sjz += rho * uz;
        }
        const R r = sr * (R)0.125;
        const R ux = (sjx * (R)0.125) / r, uy = (sjy * (R)0.125) / r, uz = (sjz * (R)0.125) / r;
        const R usq = ux * ux + uy * uy + uz * uz;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const R e = feq(i, r, ux, uy, uz, usq);
            Store<T>::store(A, (size_t)i * S + ghost[k], i, e);
            Store<T>::store(B, (size_t)i * S + ghost[k], i, e);
        }
    }
}

unsigned lgrid(size_t n) {
    size_t g = (n + kWG - 1) / kWG;
    if (g > 4096) g = 4096;
    return (unsigned)(g ? g : 1);
}

}  // namespace

template <typename T>
hipError_t launch_collide(const LevelDev& L, const BCDev& bc, const T* A, T* B, double inv_tau, hipStream_t s) {
    if (L.nactive == 0) return hipSuccess;
    unsigned g = (unsigned)((L.nactive + kBricksPerWG - 1) / kBricksPerWG);
    hipLaunchKernelGGL(k_collide<T>, dim3(g), dim3(kWG), 0, s, L, bc, A, B, (typename Store<T>::R)inv_tau);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_capture(const LevelDev& L, const BCDev& bc, const T* A, const uint32_t* slots, uint32_t ncap,
                          void* cap, hipStream_t s) {
    if (ncap == 0) return hipSuccess;
    hipLaunchKernelGGL(k_capture<T>, dim3(lgrid(ncap)), dim3(kWG), 0, s, L, bc, A, slots, ncap,
                       static_cast<typename Store<T>::R*>(cap));
    return hipGetLastError();
}

template <typename T>
hipError_t launch_fill_c2f(const uint32_t* ghost, const uint32_t* capidx, uint32_t n, const void* cap0,
                           const void* cap1, uint32_t ncap, double alpha, double fac, T* A, size_t S,
                           hipStream_t s) {
    if (n == 0) return hipSuccess;
    using R = typename Store<T>::R;
    hipLaunchKernelGGL(k_fill_c2f<T>, dim3(lgrid(n)), dim3(kWG), 0, s, ghost, capidx, n,
                       static_cast<const R*>(cap0), static_cast<const R*>(cap1), ncap, (R)alpha, (R)fac, A, S);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_fill_f2c(const uint32_t* ghost, const uint32_t* child8, uint32_t n, const LevelDev& Lf,
                           const BCDev& bc, const T* Af, double fac, T* Ac, size_t Sc, hipStream_t s) {
    if (n == 0) return hipSuccess;
    using R = typename Store<T>::R;
    hipLaunchKernelGGL(k_fill_f2c<T>, dim3(lgrid(n)), dim3(kWG), 0, s, ghost, child8, n, Lf, bc, Af, (R)fac, Ac, Sc);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_init(size_t S, double rho, const double u[3], T* A, T* B, hipStream_t s) {
    if (S == 0) return hipSuccess;
    hipLaunchKernelGGL(k_init<T>, dim3(lgrid(S)), dim3(kWG), 0, s, S, rho, u[0], u[1], u[2], A, B);
    return hipGetLastError();
}

unsigned totals_blocks(const LevelDev& L) { return lgrid((size_t)L.nactive * 64) > 1024 ? 1024 : lgrid((size_t)L.nactive * 64); }

template <typename T>
hipError_t launch_totals(const LevelDev& L, const T* A, double vol, double* partial, unsigned nblk, hipStream_t s) {
    hipLaunchKernelGGL(k_totals<T>, dim3(nblk), dim3(kWG), 0, s, L, A, vol, partial);
    return hipGetLastError();
}

unsigned force_blocks(uint32_t n) { return lgrid(n) > 1024 ? 1024 : lgrid(n); }

template <typename T>
hipError_t launch_force(const uint32_t* links, uint32_t n, size_t S, const T* A, double scale, double* partial,
                        unsigned nblk, hipStream_t s) {
    hipLaunchKernelGGL(k_force<T>, dim3(nblk), dim3(kWG), 0, s, links, n, S, A, scale, partial);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_extract(const T* A, size_t S, const int64_t* slots, size_t n, double* out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_extract<T>, dim3(lgrid(n)), dim3(kWG), 0, s, A, S, slots, n, out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_inject(T* A, T* B, size_t S, const int64_t* slots, size_t n, const double* in, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_inject<T>, dim3(lgrid(n)), dim3(kWG), 0, s, A, B, S, slots, n, in);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_vorticity(const LevelDev& L, const T* A, const int64_t* slots, size_t n, double scale, double* out,
                            hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_vorticity<T>, dim3(lgrid(n)), dim3(kWG), 0, s, L, A, slots, n, scale, out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_remap(const RemapArgs<T>& r, hipStream_t s) {
    if (r.ncopy)
        hipLaunchKernelGGL(k_remap_copy<T>, dim3(lgrid(r.ncopy)), dim3(kWG), 0, s, r.copy_src_buf, r.So, r.copy_dst,
                           r.copy_src, r.ncopy, r.A, r.B, r.S);
    if (r.nsplit)
        hipLaunchKernelGGL(k_remap_split<T>, dim3(lgrid(r.nsplit)), dim3(kWG), 0, s, r.Lc, r.bc, r.split_prev,
                           r.split_dst, r.split_src, r.nsplit, (typename Store<T>::R)r.split_fac, r.A, r.B, r.S);
    if (r.nmerge)
        hipLaunchKernelGGL(k_remap_merge<T>, dim3(lgrid(r.nmerge)), dim3(kWG), 0, s, r.Lf, r.bc, r.merge_prev,
                           r.merge_dst, r.merge_src8, r.nmerge, (typename Store<T>::R)r.merge_fac, r.A, r.B, r.S);
    if (r.nrest)
        hipLaunchKernelGGL(k_remap_rest<T>, dim3(lgrid(r.nrest)), dim3(kWG), 0, s, r.rest_dst, r.nrest, r.A, r.B, r.S);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_ghost_eq(const uint32_t* c2f_ghost, const uint32_t* c2f_cap, const uint32_t* cap_slots, uint32_t nc2f,
                           const T* Ac, size_t Sc, const uint32_t* f2c_ghost, const uint32_t* f2c_child, uint32_t nf2c,
                           const T* Af, size_t Sf, T* A, T* B, size_t S, hipStream_t s) {
    if (nc2f)
        hipLaunchKernelGGL(k_ghost_eq_c2f<T>, dim3(lgrid(nc2f)), dim3(kWG), 0, s, c2f_ghost, c2f_cap, cap_slots, nc2f,
                           Ac, Sc, A, B, S);
    if (nf2c)
        hipLaunchKernelGGL(k_ghost_eq_f2c<T>, dim3(lgrid(nf2c)), dim3(kWG), 0, s, f2c_ghost, f2c_child, nf2c, Af, Sf,
                           A, B, S);
    return hipGetLastError();
}

#define FL_INST(T)                                                                                           \
    template hipError_t launch_collide<T>(const LevelDev&, const BCDev&, const T*, T*, double, hipStream_t); \
    template hipError_t launch_capture<T>(const LevelDev&, const BCDev&, const T*, const uint32_t*, uint32_t, \
                                          void*, hipStream_t);                                               \
    template hipError_t launch_fill_c2f<T>(const uint32_t*, const uint32_t*, uint32_t, const void*, const void*, \
                                           uint32_t, double, double, T*, size_t, hipStream_t);              \
    template hipError_t launch_fill_f2c<T>(const uint32_t*, const uint32_t*, uint32_t, const LevelDev&,       \
                                           const BCDev&, const T*, double, T*, size_t, hipStream_t);        \
    template hipError_t launch_init<T>(size_t, double, const double*, T*, T*, hipStream_t);                  \
    template hipError_t launch_totals<T>(const LevelDev&, const T*, double, double*, unsigned, hipStream_t); \
    template hipError_t launch_force<T>(const uint32_t*, uint32_t, size_t, const T*, double, double*, unsigned, \
                                        hipStream_t);                                                        \
    template hipError_t launch_extract<T>(const T*, size_t, const int64_t*, size_t, double*, hipStream_t);   \
    template hipError_t launch_inject<T>(T*, T*, size_t, const int64_t*, size_t, const double*, hipStream_t); \
    template hipError_t launch_vorticity<T>(const LevelDev&, const T*, const int64_t*, size_t, double, double*,  \
                                            hipStream_t);                                                    \
    template hipError_t launch_remap<T>(const RemapArgs<T>&, hipStream_t);                                   \
    template hipError_t launch_ghost_eq<T>(const uint32_t*, const uint32_t*, const uint32_t*, uint32_t, const T*, \
                                           size_t, const uint32_t*, const uint32_t*, uint32_t, const T*, size_t, \
                                           T*, T*, size_t, hipStream_t);

FL_INST(float)
FL_INST(double)
FL_INST(_Float16)

}  // namespace lbm
}  // namespace fl
