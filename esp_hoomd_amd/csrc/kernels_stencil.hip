// PSWF charge spreading and force gather on tile-sorted particles (gfx950, wave64).
//
// One workgroup owns one spreading tile of T^3 mesh cells and the LDS box of (T+P-1)^3 cells
// its particles' stencils touch.  Particles are processed in chunks of CH: the workgroup first
// stages the chunk's charges, origins and all 3P (and, for the gather, 3P derivative) 1-D
// stencil weights in LDS in parallel, then each wave walks particles with G lanes per particle;
// a lane owns stencil columns (a, b) -- pairs of x offsets in the gather, so one ds_read2 feeds
// two columns -- and runs the P-long z loop with the z weights held in registers.
//
// Spreading accumulates in the LDS box (LDS atomics, or registers in the column-ownership form)
// and adds the box to the mesh in HBM with integer atomics: both precisions accumulate in fixed
// point (int64 cells for FP64 taps, int32 cells in LDS for FP32 taps; int64 in HBM), so the
// mesh does not depend on the particle order.  The gather reduces its per-column partial sums
// across the lanes that share a particle and adds the particle's mesh force to its output
// record with one 32-byte load and store (one thread per particle).
//
// Reference counterparts: gpu_esp_assign_particles_kernel (ESPForceComputeGPU.cu:206-272) and
// gpu_esp_interpolate_forces_kernel (ESPForceComputeGPU.cu:292-373).
#include "esp_device.h"
#include "esp_kernels.h"

#ifndef ESP_LDS_BANKS
#define ESP_LDS_BANKS 32
#endif

#include <algorithm>
#include <cstdlib>
#include <type_traits>

namespace esp
{
namespace dev
{

//! Particles staged per chunk (keeps the FP64 gather within 64 KiB of LDS at T = 8, P <= 11).
template<typename R> constexpr int chunk_size()
    {
    return sizeof(R) == 8 ? 32 : 64;
    }

//! Lanes per particle segment: a power of two in {8, 16, 32, 64} maximising lane use for
//! `items` work items per particle (ties go to the larger segment).
constexpr int choose_group(int items)
    {
    int best = 64;
    double bu = 0.0;
    const int gs[4] = {64, 32, 16, 8};
    for (int i = 0; i < 4; ++i)
        {
        const int g = gs[i];
        if (g < 16 && items > 8)
            continue;
        const int it = (items + g - 1) / g;
        const double u = double(items) / double(it * g);
        if (u > bu + 1e-9)
            {
            bu = u;
            best = g;
            }
        }
    return best;
    }

//! Stage the chunk [c0, c0 + n): charges, origins, indices and 1-D weights (and derivatives).
//! W layout: [(p * 3 + axis) * (P + 1) + k], entry k = P is a zero pad.
template<typename R, int P, bool DERIV>
__device__ __forceinline__ void stage_chunk(int c0,
                                            int n,
                                            int D,
                                            const SortedRec<R>* __restrict__ rec,
                                            const SortedIdx* __restrict__ sidx,
                                            const R* __restrict__ wc,
                                            R* __restrict__ W,
                                            R* __restrict__ DW,
                                            R* __restrict__ qs,
                                            unsigned int* __restrict__ locs,
                                            unsigned int* __restrict__ idxs)
    {
    for (int item = threadIdx.x; item < n * 3 * P; item += blockDim.x)
        {
        const int p = item / (3 * P);
        const int r = item - p * (3 * P);
        const int axis = r / P;
        const int k = r - axis * P;
        const SortedRec<R> rr = rec[c0 + p];
        const R t = axis == 0 ? rr.tx : (axis == 1 ? rr.ty : rr.tz);
        const R z = R(2) * t - R(1);
        const R* c = wc + k * (D + 1);
        R v = c[D], d = R(0);
        for (int j = D - 1; j >= 0; --j)
            {
            if (DERIV)
                d = d * z + v;
            v = v * z + c[j];
            }
        const int o = (p * 3 + axis) * (P + 1);
        W[o + k] = v;
        if (DERIV)
            DW[o + k] = R(2) * d;  // d/dt = 2 d/dz
        if (k == 0)
            {
            W[o + P] = R(0);
            if (DERIV)
                DW[o + P] = R(0);
            }
        }
    for (int p = threadIdx.x; p < n; p += blockDim.x)
        {
        qs[p] = rec[c0 + p].q;
        const SortedIdx s = sidx[c0 + p];
        locs[p] = s.local;
        idxs[p] = s.idx;
        }
    }

__device__ __forceinline__ void tile_coords(const MeshGeom& g, unsigned int tile, int& tx, int& ty, int& tz)
    {
    const int txy = g.tx * g.ty;
    tz = int(tile) / txy;
    ty = (int(tile) - tz * txy) / g.tx;
    tx = int(tile) - tz * txy - ty * g.tx;
    }

//! Walk the tile's (B^3)-cell LDS box with a block stride, calling fn(c, gx, gy, gz) with the
//! box index c and the periodic mesh coordinates of the cell.  The start and the stride are
//! split into (x, y, z) once and advanced with carries, so the loop has no integer division.
template<typename Fn>
__device__ __forceinline__ void for_box_cells(const MeshGeom& g, int tx, int ty, int tz, Fn fn)
    {
    const int B = g.B;
    const int B2 = B * B;
    const int B3 = B2 * B;
    const int t0 = int(threadIdx.x), step = int(blockDim.x);
    int lz = t0 / B2, ly = (t0 - lz * B2) / B, lx = t0 - lz * B2 - ly * B;
    const int sz = step / B2, sy = (step - sz * B2) / B, sx = step - sz * B2 - sy * B;
    const int x0 = tx * g.T, y0 = ty * g.T, z0 = tz * g.T;
    for (int c = t0; c < B3; c += step)
        {
        int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;
        while (gx >= g.nx)
            gx -= g.nx;
        while (gy >= g.ny)
            gy -= g.ny;
        while (gz >= g.nz)
            gz -= g.nz;
        fn(c, gx, gy, gz);
        lx += sx;
        ly += sy;
        lz += sz;
        if (lx >= B)
            {
            lx -= B;
            ++ly;
            }
        if (ly >= B)
            {
            ly -= B;
            ++lz;
            }
        }
    }

//! Fill the tile's LDS box from the mesh: box[c] = src[cell (gx, gy, gz)] for every box cell,
//! walked as for_box_cells does, with the global loads of NB consecutive iterations issued
//! before their LDS stores.  (A plain loop waits for each load before its store, since the
//! compiler cannot move a load above a store that may alias it: one load in flight per thread.)
template<int NB, typename L, typename T>
__device__ __forceinline__ void load_box_cells(const MeshGeom& g, int tx, int ty, int tz, const T* __restrict__ src, L* box)
    {
    const int B = g.B;
    const int B2 = B * B;
    const int B3 = B2 * B;
    const int t0 = int(threadIdx.x), step = int(blockDim.x);
    int lz = t0 / B2, ly = (t0 - lz * B2) / B, lx = t0 - lz * B2 - ly * B;
    const int sz = step / B2, sy = (step - sz * B2) / B, sx = step - sz * B2 - sy * B;
    const int x0 = tx * g.T, y0 = ty * g.T, z0 = tz * g.T;
    for (int c0 = t0; c0 < B3; c0 += NB * step)
        {
        T v[NB];
#pragma unroll
        for (int e = 0; e < NB; ++e)
            {
            if (c0 + e * step < B3)
                {
                int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;
                while (gx >= g.nx)
                    gx -= g.nx;
                while (gy >= g.ny)
                    gy -= g.ny;
                while (gz >= g.nz)
                    gz -= g.nz;
                v[e] = src[(size_t(gz) * g.ny + gy) * g.nx + gx];
                lx += sx;
                ly += sy;
                lz += sz;
                if (lx >= B)
                    {
                    lx -= B;
                    ++ly;
                    }
                if (ly >= B)
                    {
                    ly -= B;
                    ++lz;
                    }
                }
            }
#pragma unroll
        for (int e = 0; e < NB; ++e)
            if (c0 + e * step < B3)
                box[c0 + e * step] = L(v[e]);
        }
    }

// ---------------------------------------------------------------------------
// Spread
// ---------------------------------------------------------------------------
template<typename R, int P>
__global__ void __launch_bounds__(256) k_spread(MeshGeom g,
                                                int D,
                                                const unsigned int* __restrict__ tile_start,
                                                const SortedRec<R>* __restrict__ rec,
                                                const SortedIdx* __restrict__ sidx,
                                                const R* __restrict__ wcoef,
                                                typename Acc<R>::glob_t* __restrict__ mesh,
                                                const float* __restrict__ fx)
    {
    using L = typename Acc<R>::lds_t;
    constexpr int CH = chunk_size<R>();
    constexpr int ITEMS = P * P;                 // (a, b) columns per particle
    constexpr int G = choose_group(ITEMS);
    constexpr int ITER = (ITEMS + G - 1) / G;
    constexpr int PPW = WAVE / G;                // particles per wave pass

    const unsigned int tile = blockIdx.x;
    const unsigned int s0 = tile_start[tile], s1 = tile_start[tile + 1];
    if (s0 == s1)
        return;

    extern __shared__ __align__(16) unsigned char smem[];
    const int B = g.B;
    const int B2 = B * B;
    const int B3 = B2 * B;
    L* box = reinterpret_cast<L*>(smem);
    R* W = reinterpret_cast<R*>(smem + ((size_t(B3 + 1) * sizeof(L) + 15) & ~size_t(15)));
    R* wc = W + CH * 3 * (P + 1);
    R* qs = wc + P * (MAX_WINDOW_DEG + 1);
    unsigned int* locs = reinterpret_cast<unsigned int*>(qs + CH);
    unsigned int* idxs = locs + CH;

    for (int c = threadIdx.x; c <= B3; c += blockDim.x)
        box[c] = L(0);
    for (int c = threadIdx.x; c < P * (D + 1); c += blockDim.x)
        wc[c] = wcoef[c];
    const R sc = acc_scale<R>(fx);
    __syncthreads();

    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = threadIdx.x / WAVE;
    const int nw = blockDim.x / WAVE;
    const int seg = lane / G;
    const int sl = lane - seg * G;

    for (unsigned int c0 = s0; c0 < s1; c0 += CH)
        {
        const int n = int(min(unsigned(CH), s1 - c0));
        stage_chunk<R, P, false>(int(c0), n, D, rec, sidx, wc, W, nullptr, qs, locs, idxs);
        __syncthreads();
        for (int pb = wave * PPW; pb < n; pb += nw * PPW)
            {
            const int p = pb + seg;
            if (p < n)
                {
                const unsigned int loc = locs[p];
                const int ox = loc & 0xff, oy = (loc >> 8) & 0xff, oz = loc >> 16;
                const R q = qs[p];
                const R* wp = W + p * 3 * (P + 1);
                R wz[P];
#pragma unroll
                for (int c = 0; c < P; ++c)
                    wz[c] = wp[2 * (P + 1) + c];
#pragma unroll
                for (int it = 0; it < ITER; ++it)
                    {
                    const int item = sl + it * G;
                    if (item < ITEMS)
                        {
                        const int b = item / P;
                        const int a = item - b * P;
                        const R qw = q * wp[a] * wp[(P + 1) + b] * sc;
                        L* dst = box + (oz * B + oy + b) * B + ox + a;
#pragma unroll
                        for (int c = 0; c < P; ++c)
                            lds_add_scaled(dst + c * B2, qw * wz[c]);
                        }
                    }
                }
            }
        __syncthreads();
        }

    int tx, ty, tz;
    tile_coords(g, tile, tx, ty, tz);
    for (int c = threadIdx.x; c < B3; c += blockDim.x)
        {
        const L v = box[c];
        if (v != L(0))
            {
            const int lz = c / B2;
            const int ly = (c - lz * B2) / B;
            const int lx = c - lz * B2 - ly * B;
            const int gx = (tx * g.T + lx) % g.nx;
            const int gy = (ty * g.T + ly) % g.ny;
            const int gz = (tz * g.T + lz) % g.nz;
            glob_add(&mesh[(size_t(gz) * g.ny + gy) * g.nx + gx], v);
            }
        }
    }

// ---------------------------------------------------------------------------
// Gather
// ---------------------------------------------------------------------------
//! Add a particle's mesh force to the pair pass's result: one 32-byte load and one store.  Every
//! particle is gathered by exactly one thread, after the pair pass that wrote its record (or the
//! clear of the output) has completed, so no atomic is needed.
__device__ __forceinline__ void add_mesh_force(double4* __restrict__ f, double fx, double fy, double fz)
    {
    double4 v = *f;
    v.x += fx;
    v.y += fy;
    v.z += fz;
    *f = v;
    }

template<typename R, int P>
__global__ void __launch_bounds__(256) k_gather(MeshGeom g,
                                                int D,
                                                const unsigned int* __restrict__ tile_start,
                                                const SortedRec<R>* __restrict__ rec,
                                                const SortedIdx* __restrict__ sidx,
                                                const R* __restrict__ wcoef,
                                                const R* __restrict__ phi,
                                                double4* __restrict__ force)
    {
    constexpr int CH = chunk_size<R>();
    constexpr int NPAIR = (P + 1) / 2;           // x offsets handled two at a time
    constexpr int ITEMS = NPAIR * P;             // (a pair, b) columns per particle
    constexpr int G = choose_group(ITEMS);
    constexpr int ITER = (ITEMS + G - 1) / G;
    constexpr int PPW = WAVE / G;

    const unsigned int tile = blockIdx.x;
    const unsigned int s0 = tile_start[tile], s1 = tile_start[tile + 1];
    if (s0 == s1)
        return;

    extern __shared__ __align__(16) unsigned char smem[];
    const int B = g.B;
    const int B2 = B * B;
    const int B3 = B2 * B;
    R* box = reinterpret_cast<R*>(smem);
    R* W = reinterpret_cast<R*>(smem + ((size_t(B3 + 1) * sizeof(R) + 15) & ~size_t(15)));
    R* DW = W + CH * 3 * (P + 1);
    R* wc = DW + CH * 3 * (P + 1);
    R* qs = wc + P * (MAX_WINDOW_DEG + 1);
    unsigned int* locs = reinterpret_cast<unsigned int*>(qs + CH);
    unsigned int* idxs = locs + CH;

    int tx, ty, tz;
    tile_coords(g, tile, tx, ty, tz);
    load_box_cells<8>(g, tx, ty, tz, phi, box);
    if (threadIdx.x == 0)
        box[B3] = R(0);
    for (int c = threadIdx.x; c < P * (D + 1); c += blockDim.x)
        wc[c] = wcoef[c];
    __syncthreads();

    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = threadIdx.x / WAVE;
    const int nw = blockDim.x / WAVE;
    const int seg = lane / G;
    const int sl = lane - seg * G;
    const R inv_hx = R(1.0 / g.hx), inv_hy = R(1.0 / g.hy), inv_hz = R(1.0 / g.hz);

    for (unsigned int c0 = s0; c0 < s1; c0 += CH)
        {
        const int n = int(min(unsigned(CH), s1 - c0));
        stage_chunk<R, P, true>(int(c0), n, D, rec, sidx, wc, W, DW, qs, locs, idxs);
        __syncthreads();
        for (int pb = wave * PPW; pb < n; pb += nw * PPW)
            {
            const int p = pb + seg;
            R ex = R(0), ey = R(0), ez = R(0);
            if (p < n)
                {
                const unsigned int loc = locs[p];
                const int ox = loc & 0xff, oy = (loc >> 8) & 0xff, oz = loc >> 16;
                const R* wp = W + p * 3 * (P + 1);
                const R* dp = DW + p * 3 * (P + 1);
                R wz[P], dwz[P];
#pragma unroll
                for (int c = 0; c < P; ++c)
                    {
                    wz[c] = wp[2 * (P + 1) + c];
                    dwz[c] = dp[2 * (P + 1) + c];
                    }
#pragma unroll
                for (int it = 0; it < ITER; ++it)
                    {
                    const int item = sl + it * G;
                    if (item < ITEMS)
                        {
                        const int b = item / NPAIR;
                        const int a = 2 * (item - b * NPAIR);
                        const R wx0 = wp[a], wx1 = wp[a + 1];
                        const R dx0 = dp[a], dx1 = dp[a + 1];
                        const R wy = wp[(P + 1) + b], dy = dp[(P + 1) + b];
                        const R* src = box + (oz * B + oy + b) * B + ox + a;
                        R A0 = R(0), A1 = R(0), Z0 = R(0), Z1 = R(0);
#pragma unroll
                        for (int c = 0; c < P; ++c)
                            {
                            const R f0 = src[c * B2];
                            const R f1 = src[c * B2 + 1];
                            A0 += wz[c] * f0;
                            A1 += wz[c] * f1;
                            Z0 += dwz[c] * f0;
                            Z1 += dwz[c] * f1;
                            }
                        ex += wy * (dx0 * A0 + dx1 * A1);
                        ey += dy * (wx0 * A0 + wx1 * A1);
                        ez += wy * (wx0 * Z0 + wx1 * Z1);
                        }
                    }
                }
            ex = group_sum<G>(ex);
            ey = group_sum<G>(ey);
            ez = group_sum<G>(ez);
            if (p < n && sl == 0)
                {
                const double q = double(qs[p]);
                double gx = double(ex * inv_hx), gy = double(ey * inv_hy), gz = double(ez * inv_hz);
                mesh_gradient_cartesian(g, gx, gy, gz);
                add_mesh_force(force + idxs[p], q * gx, q * gy, q * gz);
                }
            }
        __syncthreads();
        }
    }

// ---------------------------------------------------------------------------
// Thread-per-particle variants: each thread owns one particle, keeps its x and z
// weights (and derivatives) in registers, evaluates the y weight per row and walks
// the P^2 rows of P contiguous LDS cells.  No cross-lane traffic at all.
// ---------------------------------------------------------------------------
//! All P taps of the window at fractional offset t (z = 2 t - 1), and with DERIV their
//! derivatives dw/dt: one Horner recurrence over the degree with the P taps side by side, so
//! the loop over the (run-time) degree carries P independent FMAs per step.  The [P][D+1]
//! coefficients are read with lane-uniform addresses.  Per tap the operations and their
//! order are those of a single-tap Horner evaluation.
template<typename R, int P, bool DERIV>
__device__ __forceinline__ void eval_taps(const R* __restrict__ wc, int D, R t, R (&w)[P], R (&dw)[P])
    {
    const R z = R(2) * t - R(1);
    const int stride = D + 1;
#pragma unroll
    for (int k = 0; k < P; ++k)
        {
        w[k] = wc[k * stride + D];
        dw[k] = R(0);
        }
    for (int j = D - 1; j >= 0; --j)
        {
#pragma unroll
        for (int k = 0; k < P; ++k)
            {
            if constexpr (DERIV)
                dw[k] = dw[k] * z + w[k];
            w[k] = w[k] * z + wc[k * stride + j];
            }
        }
    if constexpr (DERIV)
        {
#pragma unroll
        for (int k = 0; k < P; ++k)
            dw[k] = R(2) * dw[k];
        }
    }

//! LPP lanes per particle (1 or 2): with 2, lane `sub` of the pair adds rows b in
//! [sub P/2, (sub+1) P/2), so a tile's particles fill twice as many lanes (the LDS-atomic
//! stream of the tile is spread over more waves).  The fixed-point sums do not depend on it.
// ---------------------------------------------------------------------------
// Bank-balanced particle order for the thread-per-particle stencils.  The lanes of one LDS
// instruction hold different particles, each at its own box origin, and they all add the same
// offset to it; the bank an access hits is set by the origin.  Left in slot order, a tile's
// particles land on random banks and the LDS spends up to two extra cycles on conflicts per
// useful cycle.  Bucketing the particles by the bank of their origin and dealing one particle
// from every bucket in turn gives the particles that share an instruction distinct banks
// (while the buckets last).  Sums do not depend on the order: the spread's are integer, the
// gather's are per particle.
// ---------------------------------------------------------------------------
constexpr int LDS_BANKS = ESP_LDS_BANKS;
constexpr unsigned int ORDER_CAP = 512;  // larger tiles keep slot order

//! LDS bytes of the order tables (bank counters, keys, order).
constexpr size_t order_lds_bytes()
    {
    return LDS_BANKS * 4 + 2 * ORDER_CAP * 2;
    }

//! Fill order[0, n) with the tile's particle indices (0-based) in bank-balanced order; returns
//! false (and leaves the caller to use slot order) when the tile holds more than ORDER_CAP
//! particles.  `cell_words` is the LDS cell size in 4-byte words.  Every thread of the
//! workgroup must call it; it ends with a barrier.
__device__ __forceinline__ bool bank_balanced_order(unsigned int n, const SortedIdx* __restrict__ sidx, int B, int cell_words,
                                                    unsigned char* lds)
    {
    if (n > ORDER_CAP)
        return false;
    unsigned int* cnt = reinterpret_cast<unsigned int*>(lds);
    unsigned short* key = reinterpret_cast<unsigned short*>(cnt + LDS_BANKS);
    unsigned short* order = key + ORDER_CAP;
    for (int i = threadIdx.x; i < LDS_BANKS; i += blockDim.x)
        cnt[i] = 0u;
    __syncthreads();
    for (unsigned int t = threadIdx.x; t < n; t += blockDim.x)
        {
        const unsigned int loc = sidx[t].local;
        const unsigned int base = ((loc >> 16) * unsigned(B) + ((loc >> 8) & 0xffu)) * unsigned(B) + (loc & 0xffu);
        const unsigned int bank = (base * unsigned(cell_words)) % unsigned(LDS_BANKS);
        const unsigned int rank = atomicAdd(&cnt[bank], 1u);
        key[t] = static_cast<unsigned short>(rank * LDS_BANKS + bank);
        }
    __syncthreads();
    for (unsigned int t = threadIdx.x; t < n; t += blockDim.x)
        {
        const unsigned int k = key[t];
        const unsigned int rank = k / LDS_BANKS, bank = k % LDS_BANKS;
        // position: every particle of an earlier round, then this round's lower banks
        unsigned int pos = 0;
        for (int b = 0; b < LDS_BANKS; ++b)
            {
            const unsigned int c = cnt[b];
            pos += min(c, rank) + ((unsigned(b) < bank && c > rank) ? 1u : 0u);
            }
        order[pos] = static_cast<unsigned short>(t);
        }
    __syncthreads();
    return true;
    }

//! The order table written by bank_balanced_order.
__device__ __forceinline__ const unsigned short* order_table(unsigned char* lds)
    {
    return reinterpret_cast<const unsigned short*>(lds + LDS_BANKS * 4) + ORDER_CAP;
    }

template<typename R, int P, int LPP = 1, bool ORDER = false>
__device__ __forceinline__ void spread_tile_tp(unsigned int tile,
                                               const MeshGeom& g,
                                               int D,
                                               const unsigned int* __restrict__ tile_start,
                                               const SortedRec<R>* __restrict__ rec,
                                               const SortedIdx* __restrict__ sidx,
                                               const R* __restrict__ wcoef,
                                               typename Acc<R>::glob_t* __restrict__ mesh,
                                               const float* __restrict__ fx)
    {
    using L = typename Acc<R>::lds_t;
    const unsigned int s0 = tile_start[tile], s1 = tile_start[tile + 1];
    if (s0 == s1)
        return;
    extern __shared__ __align__(16) unsigned char smem[];
    const int B = g.B;
    const int B2 = B * B;
    const int B3 = B2 * B;
    L* box = reinterpret_cast<L*>(smem);
    for (int c = threadIdx.x; c <= B3; c += blockDim.x)
        box[c] = L(0);
    unsigned char* ord_lds = smem + ((size_t(B3 + 1) * sizeof(L) + 15) & ~size_t(15));
    const bool ordered = ORDER && bank_balanced_order(s1 - s0, sidx + s0, B, int(sizeof(L) / 4), ord_lds);
    const unsigned short* order = order_table(ord_lds);
    __syncthreads();

    static_assert(LPP == 1 || (LPP == 2 && P % 2 == 0), "1 or 2 lanes per particle (even P)");
    constexpr int PB = P / LPP;  // rows per lane
    const R sc = acc_scale<R>(fx);
    const unsigned int n_items = (s1 - s0) * unsigned(LPP);
    for (unsigned int t = threadIdx.x; t < n_items; t += blockDim.x)
        {
        const unsigned int p = s0 + (ordered ? unsigned(order[t / unsigned(LPP)]) : t / unsigned(LPP));
        const int sub = int(t % unsigned(LPP));
        const SortedRec<R> r = rec[p];
        const unsigned int loc = sidx[p].local;
        const int ox = loc & 0xff, oy = (loc >> 8) & 0xff, oz = loc >> 16;
        R wx[P], wy[P], wz[P], unused[P];
        eval_taps<R, P, false>(wcoef, D, r.tx, wx, unused);
        eval_taps<R, P, false>(wcoef, D, r.ty, wy, unused);
        eval_taps<R, P, false>(wcoef, D, r.tz, wz, unused);
        // this lane's rows (a select, not a run-time register index)
        R wyl[PB];
#pragma unroll
        for (int bb = 0; bb < PB; ++bb)
            wyl[bb] = (LPP == 2 && sub) ? wy[PB + bb] : wy[bb];
        L* base = box + (oz * B + oy + sub * PB) * B + ox;
        // fold the accumulator's fixed-point scale into the x weights once per particle
#pragma unroll
        for (int a = 0; a < P; ++a)
            wx[a] *= sc;
#pragma unroll
        for (int b = 0; b < PB; ++b)
            {
            const R qy = r.q * wyl[b];
            L* row = base + b * B;
#pragma unroll
            for (int c = 0; c < P; ++c)
                {
                const R v = qy * wz[c];
#pragma unroll
                for (int a = 0; a < P; ++a)
                    lds_add_scaled(row + c * B2 + a, v * wx[a]);
                }
            }
        }
    __syncthreads();

    int tx, ty, tz;
    tile_coords(g, tile, tx, ty, tz);
    for_box_cells(g, tx, ty, tz,
                  [&](int c, int gx, int gy, int gz)
                      {
                      const L v = box[c];
                      if (v != L(0))
                          glob_add(&mesh[(size_t(gz) * g.ny + gy) * g.nx + gx], v);
                      });
    }

template<typename R, int P, int LPP, bool ORDER>
__global__ void __launch_bounds__(256) k_spread_tp(MeshGeom g,
                                                   int D,
                                                   const unsigned int* __restrict__ tile_start,
                                                   const SortedRec<R>* __restrict__ rec,
                                                   const SortedIdx* __restrict__ sidx,
                                                   const R* __restrict__ wcoef,
                                                   typename Acc<R>::glob_t* __restrict__ mesh,
                                                   const float* __restrict__ fx)
    {
    spread_tile_tp<R, P, LPP, ORDER>(blockIdx.x, g, D, tile_start, rec, sidx, wcoef, mesh, fx);
    }

//! LPP lanes per particle (1 or 2): with 2, lane `sub` of the (adjacent) pair sums rows b in
//! [sub P/2, (sub+1) P/2) and the pair adds its halves with one lane exchange.
template<typename R, int P, int LPP = 1, bool ORDER = false>
__device__ __forceinline__ void gather_tile_tp(unsigned int tile,
                                               const MeshGeom& g,
                                               int D,
                                               const unsigned int* __restrict__ tile_start,
                                               const SortedRec<R>* __restrict__ rec,
                                               const SortedIdx* __restrict__ sidx,
                                               const R* __restrict__ wcoef,
                                               const R* __restrict__ phi,
                                               double4* __restrict__ force)
    {
    const unsigned int s0 = tile_start[tile], s1 = tile_start[tile + 1];
    if (s0 == s1)
        return;
    extern __shared__ __align__(16) unsigned char smem[];
    const int B = g.B;
    const int B2 = B * B;
    const int B3 = B2 * B;
    R* box = reinterpret_cast<R*>(smem);
    int tx, ty, tz;
    tile_coords(g, tile, tx, ty, tz);
    load_box_cells<8>(g, tx, ty, tz, phi, box);
    if (threadIdx.x == 0)
        box[B3] = R(0);
    unsigned char* ord_lds = smem + ((size_t(B3 + 1) * sizeof(R) + 15) & ~size_t(15));
    const bool ordered = ORDER && bank_balanced_order(s1 - s0, sidx + s0, B, int(sizeof(R) / 4), ord_lds);
    const unsigned short* order = order_table(ord_lds);
    __syncthreads();

    const R inv_hx = R(1.0 / g.hx), inv_hy = R(1.0 / g.hy), inv_hz = R(1.0 / g.hz);
    static_assert(LPP == 1 || (LPP == 2 && P % 2 == 0), "1 or 2 lanes per particle (even P)");
    constexpr int PB = P / LPP;  // rows per lane
    const unsigned int n_items = (s1 - s0) * unsigned(LPP);
    for (unsigned int t = threadIdx.x; t < n_items; t += blockDim.x)
        {
        const unsigned int p = s0 + (ordered ? unsigned(order[t / unsigned(LPP)]) : t / unsigned(LPP));
        const int sub = int(t % unsigned(LPP));
        const SortedRec<R> r = rec[p];
        const SortedIdx id = sidx[p];
        const int ox = id.local & 0xff, oy = (id.local >> 8) & 0xff, oz = id.local >> 16;
        R wx[P], dwx[P], wyf[P], dwyf[P], wz[P], dwz[P];
        eval_taps<R, P, true>(wcoef, D, r.tx, wx, dwx);
        eval_taps<R, P, true>(wcoef, D, r.ty, wyf, dwyf);
        eval_taps<R, P, true>(wcoef, D, r.tz, wz, dwz);
        // this lane's rows (a select, not a run-time register index)
        R wy[PB], dwy[PB];
#pragma unroll
        for (int bb = 0; bb < PB; ++bb)
            {
            wy[bb] = (LPP == 2 && sub) ? wyf[PB + bb] : wyf[bb];
            dwy[bb] = (LPP == 2 && sub) ? dwyf[PB + bb] : dwyf[bb];
            }
        const R* base = box + (oz * B + oy + sub * PB) * B + ox;
        R ex = R(0), ey = R(0), ez = R(0);
#pragma unroll
        for (int b = 0; b < PB; ++b)
            {
            const R* row = base + b * B;
            R Sx = R(0), Sdx = R(0), Sdz = R(0);
#pragma unroll
            for (int c = 0; c < P; ++c)
                {
                R sx = R(0), sdx = R(0);
#pragma unroll
                for (int a = 0; a < P; ++a)
                    {
                    const R f = row[c * B2 + a];
                    sx += wx[a] * f;
                    sdx += dwx[a] * f;
                    }
                Sx += wz[c] * sx;
                Sdx += wz[c] * sdx;
                Sdz += dwz[c] * sx;
                }
            ex += wy[b] * Sdx;
            ey += dwy[b] * Sx;
            ez += wy[b] * Sdz;
            }
        if constexpr (LPP == 2)
            {
            // both lanes of a pair are active together (n_items is even, pairs are lane aligned)
            ex += __shfl_xor(ex, 1, 64);
            ey += __shfl_xor(ey, 1, 64);
            ez += __shfl_xor(ez, 1, 64);
            if (sub != 0)
                continue;
            }
        const double q = double(r.q);
        double gx = double(ex * inv_hx), gy = double(ey * inv_hy), gz = double(ez * inv_hz);
        mesh_gradient_cartesian(g, gx, gy, gz);
        add_mesh_force(force + id.idx, q * gx, q * gy, q * gz);
        }
    }

template<typename R, int P, int LPP, bool ORDER>
__global__ void __launch_bounds__(256) k_gather_tp(MeshGeom g,
                                                   int D,
                                                   const unsigned int* __restrict__ tile_start,
                                                   const SortedRec<R>* __restrict__ rec,
                                                   const SortedIdx* __restrict__ sidx,
                                                   const R* __restrict__ wcoef,
                                                   const R* __restrict__ phi,
                                                   double4* __restrict__ force)
    {
    gather_tile_tp<R, P, LPP, ORDER>(blockIdx.x, g, D, tile_start, rec, sidx, wcoef, phi, force);
    }

// ---------------------------------------------------------------------------
// Matrix-core spreading (FP32): the tile's box as a GEMM.  With a 16-cell box edge (tile edge
// T = 17 - P), box[z][y][x] = sum_p wx_p[x] (q_p wy_p[y] wz_p[z]): per z-plane a 16 x 16 block
// C = Wx^T (16 x K) * Y_z (K x 16) over the tile's K particles, one v_mfma_f32_16x16x4_f32 per
// four particles and plane.  Each of the four waves owns four z-planes (four 16 x 16 fp32
// accumulators); the particles' dense, box-aligned weight vectors (zero outside the stencil)
// are staged in LDS in chunks.  Lane layout of the instruction (checked on gfx950): A[l % 16][l
// / 16], B[l / 16][l % 16], D[4 (l / 16) + r][l % 16], r = 0..3.  The sums are FP32 in the
// order of the tile's particles; the box enters the fixed-point mesh like the LDS-atomic
// variants' boxes.
// ---------------------------------------------------------------------------
typedef float mfma_f4 __attribute__((ext_vector_type(4)));

template<int P>
__global__ void __launch_bounds__(256) k_spread_mfma(MeshGeom g,
                                                     int D,
                                                     const unsigned int* __restrict__ tile_start,
                                                     const SortedRec<float>* __restrict__ rec,
                                                     const SortedIdx* __restrict__ sidx,
                                                     const float* __restrict__ wcoef,
                                                     unsigned long long* __restrict__ mesh,
                                                     const float* __restrict__ fx)
    {
    constexpr int BX = 16;    // box edge
    constexpr int CH = 64;    // particles staged per chunk
    static_assert(P >= 4 && P <= 12, "tile edge 17 - P");
    const unsigned int tile = blockIdx.x;
    const unsigned int s0 = tile_start[tile], s1 = tile_start[tile + 1];
    if (s0 == s1)
        return;
    __shared__ float wxd[CH][BX], wyd[CH][BX], qwzd[CH][BX];
    __shared__ float wc[P * (MAX_WINDOW_DEG + 1)];
    for (int c = threadIdx.x; c < P * (D + 1); c += blockDim.x)
        wc[c] = wcoef[c];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int kk = lane >> 4, col = lane & 15;
    mfma_f4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
        acc[t] = mfma_f4 {0.f, 0.f, 0.f, 0.f};

    for (unsigned int c0 = s0; c0 < s1; c0 += CH)
        {
        const int n = int(min(unsigned(CH), s1 - c0));
        for (int e = threadIdx.x; e < CH * BX; e += blockDim.x)
            {
            (&wxd[0][0])[e] = 0.f;
            (&wyd[0][0])[e] = 0.f;
            (&qwzd[0][0])[e] = 0.f;
            }
        __syncthreads();
        if (int(threadIdx.x) < n)
            {
            const int p = int(threadIdx.x);
            const SortedRec<float> r = rec[c0 + p];
            const unsigned int loc = sidx[c0 + p].local;
            const int ox = loc & 0xff, oy = (loc >> 8) & 0xff, oz = loc >> 16;
            float wx[P], wy[P], wz[P], unused[P];
            eval_taps<float, P, false>(wc, D, r.tx, wx, unused);
            eval_taps<float, P, false>(wc, D, r.ty, wy, unused);
            eval_taps<float, P, false>(wc, D, r.tz, wz, unused);
#pragma unroll
            for (int a = 0; a < P; ++a)
                {
                wxd[p][ox + a] = wx[a];
                wyd[p][oy + a] = wy[a];
                qwzd[p][oz + a] = r.q * wz[a];
                }
            }
        __syncthreads();
        for (int k0 = 0; k0 < n; k0 += 4)
            {
            const int p = k0 + kk;  // < CH: n <= CH and CH is a multiple of 4 (zero rows beyond n)
            const float a = wxd[p][col];
            const float wyv = wyd[p][col];
#pragma unroll
            for (int t = 0; t < 4; ++t)
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, wyv * qwzd[p][4 * wave + t], acc[t], 0, 0, 0);
            }
        __syncthreads();
        }

    int tx, ty, tz;
    tile_coords(g, tile, tx, ty, tz);
    const double scale = double(fx[0]);
    const int gy = (ty * g.T + col) % g.ny;
#pragma unroll
    for (int t = 0; t < 4; ++t)
        {
        const int gz = (tz * g.T + 4 * wave + t) % g.nz;
#pragma unroll
        for (int r = 0; r < 4; ++r)
            {
            const float v = acc[t][r];
            if (v != 0.f)
                {
                const int gx = (tx * g.T + 4 * kk + r) % g.nx;
                atomicAdd(&mesh[(size_t(gz) * g.ny + gy) * g.nx + gx],
                          static_cast<unsigned long long>(__double2ll_rn(double(v) * scale)));
                }
            }
        }
    }

// ---------------------------------------------------------------------------
// Column-ownership spreading: lane (cx, cy) owns the z column of the tile's
// (T+P-1)^3 box in registers.  Particles are broadcast through LDS: x and y
// weights, and the z weights already shifted to box coordinates (zero outside
// the stencil).  A lane whose column lies in a particle's xy footprint adds
// B = T+P-1 products; no LDS atomics.  FP64: every product is rounded to the fixed-point grid
// as it is added (one fma with FIX_MAGIC, summed as int64 bit patterns), so the column sums are
// exact integers whatever the order of the tile's particles.  FP32: the column sums are FP32 in
// the tile's particle order and enter the int64 mesh rounded.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void glob_add_real(unsigned long long* p, float v, double scale)
    {
    atomicAdd(p, static_cast<unsigned long long>(__double2ll_rn(double(v) * scale)));
    }

template<typename R, int P, int T>
__device__ __forceinline__ void spread_tile_own(unsigned int tile,
                                                const MeshGeom& g,
                                                int D,
                                                const unsigned int* __restrict__ tile_start,
                                                const SortedRec<R>* __restrict__ rec,
                                                const SortedIdx* __restrict__ sidx,
                                                const R* __restrict__ wcoef,
                                                typename Acc<R>::glob_t* __restrict__ mesh,
                                                const float* __restrict__ fx)
    {
    constexpr int B = T + P - 1;
    constexpr int BZ = (B + 3) & ~3;             // z weights padded to a multiple of 4
    constexpr int CH = 64;
    const unsigned int s0 = tile_start[tile], s1 = tile_start[tile + 1];
    if (s0 == s1)
        return;
    extern __shared__ __align__(16) unsigned char smem[];
    R* Wz = reinterpret_cast<R*>(smem);          // [CH][BZ] shifted z weights
    R* Wxy = Wz + CH * BZ;                        // [CH][2][P] x and y weights (times q for x)
    R* wc = Wxy + CH * 2 * P;                     // [P][D+1]
    unsigned int* org = reinterpret_cast<unsigned int*>(wc + P * (MAX_WINDOW_DEG + 1));  // [CH] ox | oy << 8
    // the chunk's records and packed origins, loaded once per particle (16-byte aligned)
    const size_t rec_off = (size_t(reinterpret_cast<unsigned char*>(org + CH) - smem) + 15) & ~size_t(15);
    SortedRec<R>* srec = reinterpret_cast<SortedRec<R>*>(smem + rec_off);  // [CH]
    unsigned int* sloc = reinterpret_cast<unsigned int*>(srec + CH);       // [CH]

    for (int c = threadIdx.x; c < P * (D + 1); c += blockDim.x)
        wc[c] = wcoef[c];

    const int cx = int(threadIdx.x) % B;
    const int cy = int(threadIdx.x) / B;
    const bool owner = int(threadIdx.x) < B * B;
    constexpr bool FIX = sizeof(R) == 8;         // FP64: int64 bit-pattern sums (see above)
    using ColT = typename std::conditional<FIX, unsigned long long, R>::type;
    ColT col[B];
#pragma unroll
    for (int z = 0; z < B; ++z)
        col[z] = ColT(0);
    unsigned int n_add = 0;                      // FP64: products added per z entry
    const R qscale = FIX ? R(fx[0]) : R(1);
    __syncthreads();

    for (unsigned int c0 = s0; c0 < s1; c0 += CH)
        {
        const int n = int(min(unsigned(CH), s1 - c0));
        for (int p = threadIdx.x; p < n; p += blockDim.x)
            {
            srec[p] = rec[c0 + p];
            sloc[p] = sidx[c0 + p].local;
            }
        __syncthreads();
        // stage: items (p, j) with j < BZ for z, BZ <= j < BZ + 2P for x and y
        constexpr int PER = BZ + 2 * P;
        for (int item = threadIdx.x; item < n * PER; item += blockDim.x)
            {
            const int p = item / PER;
            const int j = item - p * PER;
            const SortedRec<R> r = srec[p];
            const unsigned int loc = sloc[p];
            int k;
            R t;
            R scale = R(1);
            if (j < BZ)
                {
                k = j - int(loc >> 16);           // z - oz
                t = r.tz;
                }
            else if (j < BZ + P)
                {
                k = j - BZ;
                t = r.tx;
                scale = r.q * qscale;
                }
            else
                {
                k = j - BZ - P;
                t = r.ty;
                }
            R v = R(0);
            if (k >= 0 && k < P)
                {
                const R zz = R(2) * t - R(1);
                const R* c = wc + k * (D + 1);
                v = c[D];
                for (int i = D - 1; i >= 0; --i)
                    v = v * zz + c[i];
                v *= scale;
                }
            if (j < BZ)
                Wz[p * BZ + j] = v;
            else
                Wxy[p * 2 * P + (j - BZ)] = v;
            if (j == 0)
                org[p] = loc & 0xffffu;
            }
        __syncthreads();
        for (int p = 0; p < n; ++p)
            {
            const unsigned int o = org[p];
            const int dx = cx - int(o & 0xff);
            const int dy = cy - int(o >> 8);
            if (owner && dx >= 0 && dx < P && dy >= 0 && dy < P)
                {
                const R w = Wxy[p * 2 * P + dx] * Wxy[p * 2 * P + P + dy];
                const R* wz = Wz + p * BZ;
                if constexpr (FIX)
                    {
#pragma unroll
                    for (int z = 0; z < B; ++z)
                        col[z] += static_cast<unsigned long long>(__double_as_longlong(fma(w, wz[z], FIX_MAGIC)));
                    ++n_add;
                    }
                else
                    {
#pragma unroll
                    for (int z = 0; z < B; ++z)
                        col[z] += w * wz[z];
                    }
                }
            }
        __syncthreads();
        }

    if (owner)
        {
        int tx, ty, tz;
        tile_coords(g, tile, tx, ty, tz);
        const int gx = (tx * T + cx) % g.nx;
        const int gy = (ty * T + cy) % g.ny;
        if constexpr (FIX)
            {
            // remove the magic constant of every product added (modulo 2^64, like the sums)
            const unsigned long long bias = static_cast<unsigned long long>(n_add) * static_cast<unsigned long long>(FIX_MAGIC_BITS);
#pragma unroll
            for (int z = 0; z < B; ++z)
                {
                const unsigned long long v = col[z] - bias;
                if (v != 0ull)
                    atomicAdd(&mesh[(size_t((tz * T + z) % g.nz) * g.ny + gy) * g.nx + gx], v);
                }
            }
        else
            {
            const double scale = double(fx[0]);
#pragma unroll
            for (int z = 0; z < B; ++z)
                if (col[z] != R(0))
                    {
                    const int gz = (tz * T + z) % g.nz;
                    glob_add_real(&mesh[(size_t(gz) * g.ny + gy) * g.nx + gx], col[z], scale);
                    }
            }
        }
    }

template<typename R, int P, int T>
__global__ void __launch_bounds__(1024) k_spread_own(MeshGeom g,
                                                     int D,
                                                     const unsigned int* __restrict__ tile_start,
                                                     const SortedRec<R>* __restrict__ rec,
                                                     const SortedIdx* __restrict__ sidx,
                                                     const R* __restrict__ wcoef,
                                                     typename Acc<R>::glob_t* __restrict__ mesh,
                                                     const float* __restrict__ fx)
    {
    spread_tile_own<R, P, T>(blockIdx.x, g, D, tile_start, rec, sidx, wcoef, mesh, fx);
    }

// ---------------------------------------------------------------------------
// Batched spreading and gathering: items are (system, tile); the stencil order is a
// per-system runtime value dispatched to the compiled orders.
// ---------------------------------------------------------------------------
//! Lanes per particle of the two-lane thread-per-particle variant: 2 for even orders (half the
//! rows each), 1 for odd ones.
constexpr int tp_pair_lanes(int P)
    {
    return P % 2 == 0 ? 2 : 1;
    }

#define ESP_BATCH_P_CASES(X) X(4) X(6) X(8) X(10) X(12)

template<typename R, bool ORDER>
__global__ void __launch_bounds__(256) kb_spread(const kernel::BatchSys* __restrict__ ds, kernel::BatchList l)
    {
    for (unsigned int it = blockIdx.x; it < l.n; it += gridDim.x)
        {
        const kernel::BatchItem w = l.items[it];
        const kernel::BatchSys& d = ds[w.sys];
        const auto* rec = reinterpret_cast<const SortedRec<R>*>(d.rec);
        const auto* wc = reinterpret_cast<const R*>(d.wcoef);
        auto* mesh = reinterpret_cast<typename Acc<R>::glob_t*>(d.mesh);
        switch (d.P)
            {
#define ESP_CASE(PP)                                                                 \
    case PP:                                                                         \
        spread_tile_tp<R, PP, 1, ORDER>(w.local, d.g, d.D, d.tile_start, rec, d.sidx, wc, mesh, d.fx); \
        break;
            ESP_BATCH_P_CASES(ESP_CASE)
#undef ESP_CASE
            default: break;
            }
        __syncthreads();
        }
    }

template<typename R, bool ORDER>
__global__ void __launch_bounds__(256) kb_gather(const kernel::BatchSys* __restrict__ ds, kernel::BatchList l)
    {
    for (unsigned int it = blockIdx.x; it < l.n; it += gridDim.x)
        {
        const kernel::BatchItem w = l.items[it];
        const kernel::BatchSys& d = ds[w.sys];
        const auto* rec = reinterpret_cast<const SortedRec<R>*>(d.rec);
        const auto* wc = reinterpret_cast<const R*>(d.wcoef);
        const auto* phi = reinterpret_cast<const R*>(d.phi);
        auto* force = reinterpret_cast<double4*>(d.force);
        switch (d.P)
            {
#define ESP_CASE(PP)                                                                          \
    case PP:                                                                                  \
        gather_tile_tp<R, PP, 1, ORDER>(w.local, d.g, d.D, d.tile_start, rec, d.sidx, wc, phi, force); \
        break;
            ESP_BATCH_P_CASES(ESP_CASE)
#undef ESP_CASE
            default: break;
            }
        __syncthreads();
        }
    }

    } // namespace dev

namespace kernel
{
using namespace esp::dev;

size_t own_lds_bytes(int P, bool fp64)
    {
    const size_t el = fp64 ? 8 : 4;
    const int B = 8 + P - 1;  // largest box edge (T = 8)
    const size_t BZ = size_t((B + 3) & ~3);
    const size_t base = 64 * BZ * el + 64 * 2 * size_t(P) * el + size_t(P) * (MAX_WINDOW_DEG + 1) * el + 64 * 4;
    // + the chunk's records (16-byte aligned) and packed origins
    return ((base + 15) & ~size_t(15)) + 64 * 4 * el + 64 * 4;
    }

template<typename R, int P>
static void launch_spread_own(const MeshGeom& g, int D, unsigned int n_tiles, const unsigned int* start, const void* rec,
                              const SortedIdx* sidx, const void* wcoef, void* mesh, const float* fx, hipStream_t s)
    {
    const size_t lds = own_lds_bytes(P, sizeof(R) == 8);
    const unsigned int threads = unsigned((g.B * g.B + 63) / 64 * 64);
    const auto* rr = reinterpret_cast<const SortedRec<R>*>(rec);
    const auto* wc = reinterpret_cast<const R*>(wcoef);
    auto* m = reinterpret_cast<typename Acc<R>::glob_t*>(mesh);
    if (g.T == 8)
        hipLaunchKernelGGL((k_spread_own<R, P, 8>), dim3(n_tiles), dim3(threads), lds, s, g, D, start, rr, sidx, wc, m, fx);
    else if (g.T == 4)
        hipLaunchKernelGGL((k_spread_own<R, P, 4>), dim3(n_tiles), dim3(threads), lds, s, g, D, start, rr, sidx, wc, m, fx);
    else
        hipLaunchKernelGGL((k_spread_own<R, P, 2>), dim3(n_tiles), dim3(threads), lds, s, g, D, start, rr, sidx, wc, m, fx);
    }

//! Stencil kernel variant: 0 = wave segments per particle, 1 = thread per particle,
//! 2 = column ownership (spread only), 3 = thread per particle with two lanes per particle
//! (even orders; half the rows each), 4 = matrix cores (FP32 spread with a 16-cell box, tile
//! edge 17 - P; other precisions and the gather use their defaults).
//! Defaults (measured on MI355X at N = 1e6, P = 10): both precisions spread with two lanes per
//! particle (fixed-point sums: the same result as one lane, so as the batched kernels; FP64
//! 0.49 ms against 0.61 ms for column ownership with its per-product int64 rounding at 180^3)
//! and gather with one, as the batched kernels do (their 12^3 tiles already fill the
//! workgroup; FP64 0.433 ms against 0.427 ms with two lanes), so a system's forces are
//! bit-identical alone and in a batch in both precisions.  The environment variables
//! ESP_SPREAD_VARIANT / ESP_GATHER_VARIANT override them.
static int stencil_variant(int which, bool fp64)
    {
    static int v[2] = {-2, -2};
    if (v[which] == -2)
        {
        const char* e = getenv(which == 0 ? "ESP_SPREAD_VARIANT" : "ESP_GATHER_VARIANT");
        v[which] = e ? atoi(e) : -1;
        }
    if (v[which] >= 0 && !(v[which] == 4 && (fp64 || which != 0)))
        return v[which];
    (void)fp64;
    return which == 0 ? 3 : 1;
    }

int mfma_spread_tile(int P, bool fp64)
    {
    return (!fp64 && stencil_variant(0, false) == 4 && P >= 4 && P <= 12) ? 17 - P : 0;
    }

size_t tp_lds_bytes(const MeshGeom& g, int P, bool fp64);
size_t spread_lds_bytes(const MeshGeom& g, int P, int D, bool fp64);
size_t gather_lds_bytes(const MeshGeom& g, int P, int D, bool fp64);

//! Whether the spreading/gathering kernels selected for this precision can run on tiles of
//! edge T at stencil order P: their LDS box must fit 64 KiB, and column-ownership spreading
//! has instantiations for T = 2, 4, 8 only.  `batched` asks for the batched kernels (which
//! always use the thread-per-particle forms).
bool stencil_tile_ok(int T, int P, bool fp64, bool batched)
    {
    MeshGeom g {};
    g.T = T;
    g.B = T + P - 1;
    const int sv = batched ? 1 : stencil_variant(0, fp64);
    const int gv = batched ? 1 : stencil_variant(1, fp64);
    size_t lds = 0;
    if (sv == 2)
        {
        if (T != 2 && T != 4 && T != 8)  // the instantiated box edges
            return false;
        lds = own_lds_bytes(P, fp64);
        }
    else if (sv == 4)
        {
        if (T != 17 - P)  // 16-cell box
            return false;
        lds = 3 * 64 * 16 * 4 + size_t(P) * (MAX_WINDOW_DEG + 1) * 4;
        }
    else
        lds = (sv == 1 || sv == 3) ? tp_lds_bytes(g, P, fp64) : spread_lds_bytes(g, P, 0, fp64);
    lds = std::max(lds, (gv == 1 || gv == 3) ? tp_lds_bytes(g, P, fp64) : gather_lds_bytes(g, P, 0, fp64));
    return lds <= 64 * 1024;
    }

size_t tp_lds_bytes(const MeshGeom& g, int P, bool fp64)
    {
    const size_t B3 = size_t(g.B) * g.B * g.B;
    const size_t el = fp64 ? 8 : 4;
    (void)P;
    return (((B3 + 1) * el + 15) & ~size_t(15)) + order_lds_bytes();
    }

size_t spread_lds_bytes(const MeshGeom& g, int P, int D, bool fp64)
    {
    (void)D;
    const size_t B3 = size_t(g.B) * g.B * g.B;
    const size_t el = fp64 ? 8 : 4;
    const size_t CH = fp64 ? chunk_size<double>() : chunk_size<float>();
    const size_t box = ((B3 + 1) * el + 15) & ~size_t(15);
    const size_t w = size_t(CH) * 3 * (P + 1) * el;
    const size_t wc = size_t(P) * (MAX_WINDOW_DEG + 1) * el;
    return box + w + wc + CH * el + 2 * CH * 4;
    }

size_t gather_lds_bytes(const MeshGeom& g, int P, int D, bool fp64)
    {
    const size_t el = fp64 ? 8 : 4;
    const size_t CH = fp64 ? chunk_size<double>() : chunk_size<float>();
    return spread_lds_bytes(g, P, D, fp64) + CH * 3 * (P + 1) * el;
    }

//! Bank-balanced particle order (bank_balanced_order) in the spread (which = 0) and the gather
//! (which = 1).  Default: the gather only.  Measured on MI355X at 1e6 particles, 160^3, P = 10
//! (FP32): it cuts the gather's bank-conflict cycles per LDS cycle from 1.9 to 1.2 and the
//! kernel from 229 to 213 us; the spread's conflicts fall from 1.7 to 1.2 but the kernel does not
//! speed up (318 against 309 us: it is bound by the LDS atomics' own rate).  ESP_BANK_ORDER
//! (bit 0 spread, bit 1 gather) overrides it.
static bool bank_order(int which)
    {
    static int mask = -1;
    if (mask < 0)
        {
        const char* e = getenv("ESP_BANK_ORDER");
        mask = e ? atoi(e) : 2;
        }
    return (mask >> which) & 1;
    }

template<typename R, int P, int LPP>
static void launch_spread_tp(bool order, unsigned int n_tiles, size_t lds, hipStream_t s, const MeshGeom& g, int D,
                             const unsigned int* start, const SortedRec<R>* rec, const SortedIdx* sidx, const R* wc,
                             typename Acc<R>::glob_t* mesh, const float* fx)
    {
    if (order)
        hipLaunchKernelGGL((k_spread_tp<R, P, LPP, true>), dim3(n_tiles), dim3(256), lds, s, g, D, start, rec, sidx, wc, mesh, fx);
    else
        hipLaunchKernelGGL((k_spread_tp<R, P, LPP, false>), dim3(n_tiles), dim3(256), lds, s, g, D, start, rec, sidx, wc, mesh, fx);
    }

template<typename R, int P, int LPP>
static void launch_gather_tp(bool order, unsigned int n_tiles, size_t lds, hipStream_t s, const MeshGeom& g, int D,
                             const unsigned int* start, const SortedRec<R>* rec, const SortedIdx* sidx, const R* wc, const R* phi,
                             double4* force)
    {
    if (order)
        hipLaunchKernelGGL((k_gather_tp<R, P, LPP, true>), dim3(n_tiles), dim3(256), lds, s, g, D, start, rec, sidx, wc, phi, force);
    else
        hipLaunchKernelGGL((k_gather_tp<R, P, LPP, false>), dim3(n_tiles), dim3(256), lds, s, g, D, start, rec, sidx, wc, phi, force);
    }

#define ESP_P_CASES(X) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16)

template<typename R>
hipError_t spread(const MeshGeom& g, int P, int D, unsigned int n_tiles, const unsigned int* start, const void* rec,
                  const SortedIdx* sidx, const void* wcoef, void* mesh, const float* fx, hipStream_t s)
    {
    const int variant = stencil_variant(0, sizeof(R) == 8);
    const bool tp = variant == 1 || variant == 3;
    const size_t lds = tp ? tp_lds_bytes(g, P, sizeof(R) == 8) : spread_lds_bytes(g, P, D, sizeof(R) == 8);
    switch (P)
        {
#define ESP_CASE(PP)                                                                                                         \
    case PP:                                                                                                                 \
        if (variant == 4 && sizeof(R) == 4 && PP >= 4 && PP <= 12)                                                          \
            hipLaunchKernelGGL((k_spread_mfma<(PP >= 4 && PP <= 12) ? PP : 4>), dim3(n_tiles), dim3(256), 0, s, g, D, start,  \
                               reinterpret_cast<const SortedRec<float>*>(rec), sidx, reinterpret_cast<const float*>(wcoef), \
                               reinterpret_cast<unsigned long long*>(mesh), fx);                                           \
        else if (variant == 2)                                                                                               \
            launch_spread_own<R, PP>(g, D, n_tiles, start, rec, sidx, wcoef, mesh, fx, s);                                  \
        else if (variant == 3)                                                                                               \
            launch_spread_tp<R, PP, tp_pair_lanes(PP)>(bank_order(0), n_tiles, lds, s, g, D, start,                              \
                               reinterpret_cast<const SortedRec<R>*>(rec), sidx, reinterpret_cast<const R*>(wcoef),          \
                               reinterpret_cast<typename Acc<R>::glob_t*>(mesh), fx);                                        \
        else if (tp)                                                                                                         \
            launch_spread_tp<R, PP, 1>(bank_order(0), n_tiles, lds, s, g, D, start,                                             \
                               reinterpret_cast<const SortedRec<R>*>(rec), sidx, reinterpret_cast<const R*>(wcoef),          \
                               reinterpret_cast<typename Acc<R>::glob_t*>(mesh), fx);                                        \
        else                                                                                                                 \
            hipLaunchKernelGGL((k_spread<R, PP>), dim3(n_tiles), dim3(256), lds, s, g, D, start,                            \
                               reinterpret_cast<const SortedRec<R>*>(rec), sidx, reinterpret_cast<const R*>(wcoef),          \
                               reinterpret_cast<typename Acc<R>::glob_t*>(mesh), fx);                                        \
        break;
        ESP_P_CASES(ESP_CASE)
#undef ESP_CASE
        default: return hipErrorInvalidValue;
        }
    return hipGetLastError();
    }

template<typename R>
hipError_t gather(const MeshGeom& g, int P, int D, unsigned int n_tiles, const unsigned int* start, const void* rec,
                  const SortedIdx* sidx, const void* wcoef, const void* phi, double* force, hipStream_t s)
    {
    const int variant = stencil_variant(1, sizeof(R) == 8);
    const bool tp = variant == 1 || variant == 3;
    const size_t lds = tp ? tp_lds_bytes(g, P, sizeof(R) == 8) : gather_lds_bytes(g, P, D, sizeof(R) == 8);
    switch (P)
        {
#define ESP_CASE(PP)                                                                                                         \
    case PP:                                                                                                                 \
        if (variant == 3)                                                                                                    \
            launch_gather_tp<R, PP, tp_pair_lanes(PP)>(bank_order(1), n_tiles, lds, s, g, D, start,                              \
                               reinterpret_cast<const SortedRec<R>*>(rec), sidx, reinterpret_cast<const R*>(wcoef),          \
                               reinterpret_cast<const R*>(phi), reinterpret_cast<double4*>(force));                          \
        else if (tp)                                                                                                         \
            launch_gather_tp<R, PP, 1>(bank_order(1), n_tiles, lds, s, g, D, start,                                             \
                               reinterpret_cast<const SortedRec<R>*>(rec), sidx, reinterpret_cast<const R*>(wcoef),          \
                               reinterpret_cast<const R*>(phi), reinterpret_cast<double4*>(force));                          \
        else                                                                                                                 \
            hipLaunchKernelGGL((k_gather<R, PP>), dim3(n_tiles), dim3(256), lds, s, g, D, start,                            \
                               reinterpret_cast<const SortedRec<R>*>(rec), sidx, reinterpret_cast<const R*>(wcoef),          \
                               reinterpret_cast<const R*>(phi), reinterpret_cast<double4*>(force));                          \
        break;
        ESP_P_CASES(ESP_CASE)
#undef ESP_CASE
        default: return hipErrorInvalidValue;
        }
    return hipGetLastError();
    }

template hipError_t spread<double>(const MeshGeom&, int, int, unsigned int, const unsigned int*, const void*, const SortedIdx*,
                                   const void*, void*, const float*, hipStream_t);
template hipError_t spread<float>(const MeshGeom&, int, int, unsigned int, const unsigned int*, const void*, const SortedIdx*,
                                  const void*, void*, const float*, hipStream_t);
template hipError_t gather<double>(const MeshGeom&, int, int, unsigned int, const unsigned int*, const void*, const SortedIdx*,
                                   const void*, const void*, double*, hipStream_t);
template hipError_t gather<float>(const MeshGeom&, int, int, unsigned int, const unsigned int*, const void*, const SortedIdx*,
                                  const void*, const void*, double*, hipStream_t);

bool batch_order_supported(int P)
    {
    switch (P)
        {
#define ESP_CASE(PP) case PP:
        ESP_BATCH_P_CASES(ESP_CASE)
#undef ESP_CASE
        return true;
        default: return false;
        }
    }

template<typename R> hipError_t batch_spread(const BatchSys* d, BatchList l, size_t lds, hipStream_t s)
    {
    if (l.n == 0)
        return hipSuccess;
    if (bank_order(0))
        hipLaunchKernelGGL((kb_spread<R, true>), dim3(l.grid), dim3(256), lds, s, d, l);
    else
        hipLaunchKernelGGL((kb_spread<R, false>), dim3(l.grid), dim3(256), lds, s, d, l);
    return hipGetLastError();
    }

template<typename R> hipError_t batch_gather(const BatchSys* d, BatchList l, size_t lds, hipStream_t s)
    {
    if (l.n == 0)
        return hipSuccess;
    if (bank_order(1))
        hipLaunchKernelGGL((kb_gather<R, true>), dim3(l.grid), dim3(256), lds, s, d, l);
    else
        hipLaunchKernelGGL((kb_gather<R, false>), dim3(l.grid), dim3(256), lds, s, d, l);
    return hipGetLastError();
    }

template hipError_t batch_spread<double>(const BatchSys*, BatchList, size_t, hipStream_t);
template hipError_t batch_spread<float>(const BatchSys*, BatchList, size_t, hipStream_t);
template hipError_t batch_gather<double>(const BatchSys*, BatchList, size_t, hipStream_t);
template hipError_t batch_gather<float>(const BatchSys*, BatchList, size_t, hipStream_t);

    } // namespace kernel
    } // namespace esp
