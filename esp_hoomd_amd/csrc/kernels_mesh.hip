// ESP mesh kernels for gfx950: tile binning, LDS-tiled PSWF spreading, Stockham FFT passes with
// the influence multiply fused into the z pass, the mirrored gather, influence tabulation and
// the FP64 energy/virial reduction.
//
// Reference counterparts: gpu_esp_assign_particles_kernel (ESPForceComputeGPU.cu:206-272),
// hipfftExecC2C (ESPForceComputeGPU.cc:254-285), gpu_esp_apply_influence_kernel (.cu:277-287),
// gpu_esp_interpolate_forces_kernel (.cu:292-373), gpu_esp_build_influence_kernel (.cu:157-201).
#include "esp_device.h"
#include "esp_kernels.h"

namespace esp
{
namespace dev
{

//! Read a fixed-point mesh cell as a real (times 1 / S) and clear it for the next step.
template<typename R> __device__ __forceinline__ R mesh_load_clear(unsigned long long* p, double inv_scale)
    {
    const long long v = static_cast<long long>(*p);
    *p = 0ull;
    return static_cast<R>(static_cast<double>(v) * inv_scale);
    }

//! The step's fixed-point scale S for the mesh accumulation: fx[0] = S, fx[1] = 1 / S, fx[2] = 1
//! when a population bound lowered S.  S is a power of two, so both are exact in FP32.
//!  * FP32 taps into int32 LDS cells (rule.wide = 0): S = 2^23 / 2^ceil(log2 max|q|) (clamped to
//!    [1, 2^40]) keeps the quantisation step relative to the charges (2^-23 of the largest), and
//!    is halved while the most populated tile could overflow an int32 LDS cell: a cell of a
//!    tile box sums at most `cmax` contributions, each |q w_x w_y w_z| S <= max|q| w3max S
//!    rounded.
//!  * FP64 taps into int64 cells (rule.wide = 1): S = 2^50 / 2^ceil(log2 max|q|) (clamped to
//!    [2^-100, 2^100]), a step of 2^-50 of the largest charge (FP64 keeps 2^-52 of a value);
//!    halved while a tap could reach 2^51 (the limit of round_to_ll) or while the cells of the
//!    `cover` tile boxes that overlap one mesh cell, `cmax` contributions each, could overflow
//!    int64.
//! Without halving, S depends only on the charges, so every tiling gives the same sums.
__device__ __forceinline__ void fixed_point_scale(unsigned int cmax, double qmax, const FixedPointRule& rule, float* fx)
    {
    const double qm = qmax > 0.0 ? qmax : 1.0;
    int e = 0;
    const double m = frexp(qm, &e);  // qm = m 2^e, m in [0.5, 1)
    const int ceil_log2 = (m == 0.5) ? e - 1 : e;
    const double tap = qm * double(rule.w3max);
    float lowered = 0.0f;
    double S;
    if (rule.wide)
        {
        int k = 50 - ceil_log2;
        k = k < -100 ? -100 : (k > 100 ? 100 : k);
        S = ldexp(1.0, k);
        const double tap_limit = 2251799813685248.0;                  // 2^51
        const double sum_limit = 0.9 * 9223372036854775808.0;         // 0.9 * 2^63
        const double n = double(rule.cover) * double(cmax);
        while (k > -120 && (tap * S >= tap_limit || n * (tap * S + 0.5) >= sum_limit))
            {
            S *= 0.5;
            --k;
            lowered = 1.0f;
            }
        }
    else
        {
        int k = 23 - ceil_log2;
        k = k < 0 ? 0 : (k > 40 ? 40 : k);
        S = ldexp(1.0, k);
        const double limit = 0.9 * 2147483648.0;
        while (S > 1.0 && double(cmax) * (tap * S + 0.5) >= limit)
            {
            S *= 0.5;
            lowered = 1.0f;
            }
        }
    fx[0] = float(S);
    fx[1] = float(1.0 / S);
    fx[2] = lowered;
    }

//! Stencil origin o = ceil(g - P/2) (wrapped into [0, n)) and offset t = o - (g - P/2) in [0, 1)
//! for one axis; g = makeFraction(r) * n, from the unwrapped fraction u of box_frac.
__device__ __forceinline__ int stencil_origin(double u, int n, int P, double& t)
    {
    double frac = u;
    frac -= floor(frac);
    const double gg = frac * double(n) - 0.5 * double(P);
    const double o = ceil(gg);
    t = o - gg;
    int oi = int(o) % n;
    if (oi < 0)
        oi += n;
    return oi;
    }

// ---------------------------------------------------------------------------
// Binning into spreading tiles (counting sort: count, scan, scatter)
// ---------------------------------------------------------------------------
//! Bin group member i into its spreading tile; returns its charge.
__device__ __forceinline__ double bin_one(unsigned int i,
                                          const unsigned int* __restrict__ group,
                                          const double4* __restrict__ postype,
                                          const double* __restrict__ charge,
                                          const MeshGeom& g,
                                          int P,
                                          unsigned int* __restrict__ tile_count,
                                          unsigned int* __restrict__ key,
                                          unsigned int* __restrict__ slot)
    {
    const unsigned int p = group ? group[i] : i;
    const double4 r = postype[p];
    const double q = charge[p];
    unsigned int k = 0xffffffffu;
    if (q != 0.0)
        {
        double t;
        const Frac3 u = box_frac(r.x, r.y, r.z, g);
        const int ox = stencil_origin(u.x, g.nx, P, t);
        const int oy = stencil_origin(u.y, g.ny, P, t);
        const int oz = stencil_origin(u.z, g.nz, P, t);
        k = unsigned((oz / g.T) * g.ty + oy / g.T) * unsigned(g.tx) + unsigned(ox / g.T);
        slot[i] = atomicAdd(&tile_count[k], 1u);
        }
    key[i] = k;
    return q;
    }

__global__ void __launch_bounds__(256) k_bin(unsigned int n,
                                             const unsigned int* __restrict__ group,
                                             const double4* __restrict__ postype,
                                             const double* __restrict__ charge,
                                             MeshGeom g,
                                             int P,
                                             unsigned int* __restrict__ tile_count,
                                             unsigned int* __restrict__ key,
                                             unsigned int* __restrict__ slot,
                                             double* __restrict__ qpart)
    {
    const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    double q = 0.0;
    if (i < n)
        q = bin_one(i, group, postype, charge, g, P, tile_count, key, slot);
    // block partial sums of q and q^2 (deterministic order) and the largest |q|
    __shared__ double sq[8], sq2[8], sqm[8];
    double a = wave_sum(q), b = wave_sum(q * q), c = wave_max(fabs(q));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        {
        sq[wave] = a;
        sq2[wave] = b;
        sqm[wave] = c;
        }
    __syncthreads();
    if (threadIdx.x == 0)
        {
        double s = 0, s2 = 0, sm = 0;
        for (int w = 0; w < int(blockDim.x >> 6); ++w)
            {
            s += sq[w];
            s2 += sq2[w];
            sm = fmax(sm, sqm[w]);
            }
        qpart[3 * blockIdx.x] = s;
        qpart[3 * blockIdx.x + 1] = s2;
        qpart[3 * blockIdx.x + 2] = sm;
        }
    }

//! Exclusive scan of tile counts by one 1024-thread workgroup; clears the counts for the next
//! step.  Rounds of 16 consecutive counts per thread: the 16 loads of a thread are independent
//! and in flight together, the thread sums are scanned with wave shuffles, then across the 16
//! waves.
__device__ __forceinline__ void scan_one(unsigned int n_tiles, unsigned int* __restrict__ count, unsigned int* __restrict__ start,
                                         int n_q, const double* __restrict__ qpart, FixedPointRule rule, float* __restrict__ fx)
    {
    constexpr int CH = 16;
    __shared__ unsigned int wsum[16];
    __shared__ unsigned int wmax[16];
    __shared__ double wq[16];
    const unsigned int tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    unsigned int carry = 0;
    unsigned int cmax = 0;
    for (unsigned int base = 0; base < n_tiles; base += 1024u * CH)
        {
        const unsigned int b0 = base + tid * CH;
        unsigned int v[CH];
        unsigned int s = 0;
#pragma unroll
        for (int k = 0; k < CH; ++k)
            {
            const unsigned int i = b0 + unsigned(k);
            v[k] = i < n_tiles ? count[i] : 0u;
            }
#pragma unroll
        for (int k = 0; k < CH; ++k)
            {
            s += v[k];
            cmax = max(cmax, v[k]);
            }
        // inclusive scan of the thread sums within the wave
        unsigned int x = s;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1)
            {
            const unsigned int y = __shfl_up(x, off);
            if (int(lane) >= off)
                x += y;
            }
        if (lane == 63u)
            wsum[wave] = x;
        __syncthreads();
        if (wave == 0)
            {
            unsigned int w = lane < 16u ? wsum[lane] : 0u;
#pragma unroll
            for (int off = 1; off < 16; off <<= 1)
                {
                const unsigned int y = __shfl_up(w, off);
                if (int(lane) >= off)
                    w += y;
                }
            if (lane < 16u)
                wsum[lane] = w;
            }
        __syncthreads();
        unsigned int run = carry + (wave ? wsum[wave - 1] : 0u) + x - s;
#pragma unroll
        for (int k = 0; k < CH; ++k)
            {
            const unsigned int i = b0 + unsigned(k);
            if (i < n_tiles)
                {
                start[i] = run;
                count[i] = 0u;
                }
            run += v[k];
            }
        carry += wsum[15];
        __syncthreads();
        }
    if (tid == 0)
        start[n_tiles] = carry;
    if (fx)
        {
        // the most populated tile and the largest |q| (from the binning partials)
        double qm = 0.0;
        for (int i = int(tid); i < n_q; i += int(blockDim.x))
            qm = fmax(qm, qpart[3 * i + 2]);
        cmax = wave_max(cmax);
        qm = wave_max(qm);
        if (lane == 0)
            {
            wmax[wave] = cmax;
            wq[wave] = qm;
            }
        __syncthreads();
        if (tid == 0)
            {
            for (unsigned int w = 1; w < blockDim.x / 64u; ++w)
                {
                cmax = max(cmax, wmax[w]);
                qm = fmax(qm, wq[w]);
                }
            fixed_point_scale(cmax, qm, rule, fx);
            }
        }
    }

__global__ void __launch_bounds__(1024) k_scan(unsigned int n_tiles, unsigned int* __restrict__ count, unsigned int* __restrict__ start,
                                              int n_q, const double* __restrict__ qpart, FixedPointRule rule, float* __restrict__ fx)
    {
    scan_one(n_tiles, count, start, n_q, qpart, rule, fx);
    }

//! Write group member i's record into its tile-sorted slot.
template<typename R>
__device__ __forceinline__ void scatter_one(unsigned int i,
                                            const unsigned int* __restrict__ group,
                                            const double4* __restrict__ postype,
                                            const double* __restrict__ charge,
                                            const MeshGeom& g,
                                            int P,
                                            const unsigned int* __restrict__ key,
                                            const unsigned int* __restrict__ slot,
                                            const unsigned int* __restrict__ start,
                                            SortedRec<R>* __restrict__ rec,
                                            SortedIdx* __restrict__ sidx)
    {
    const unsigned int k = key[i];
    if (k == 0xffffffffu)
        return;
    const unsigned int p = group ? group[i] : i;
    const double4 r = postype[p];
    double tx, ty, tz;
    const Frac3 u = box_frac(r.x, r.y, r.z, g);
    const int ox = stencil_origin(u.x, g.nx, P, tx);
    const int oy = stencil_origin(u.y, g.ny, P, ty);
    const int oz = stencil_origin(u.z, g.nz, P, tz);
    const unsigned int pos = start[k] + slot[i];
    SortedRec<R> o;
    o.tx = R(tx);
    o.ty = R(ty);
    o.tz = R(tz);
    o.q = R(charge[p]);
    rec[pos] = o;
    SortedIdx s;
    s.local = unsigned(ox % g.T) | (unsigned(oy % g.T) << 8) | (unsigned(oz % g.T) << 16);
    s.idx = p;
    sidx[pos] = s;
    }

template<typename R>
__global__ void __launch_bounds__(256) k_scatter(unsigned int n,
                                                 const unsigned int* __restrict__ group,
                                                 const double4* __restrict__ postype,
                                                 const double* __restrict__ charge,
                                                 MeshGeom g,
                                                 int P,
                                                 const unsigned int* __restrict__ key,
                                                 const unsigned int* __restrict__ slot,
                                                 const unsigned int* __restrict__ start,
                                                 SortedRec<R>* __restrict__ rec,
                                                 SortedIdx* __restrict__ sidx)
    {
    const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        scatter_one<R>(i, group, postype, charge, g, P, key, slot, start, rec, sidx);
    }

// ---------------------------------------------------------------------------
// FFT passes.  Layout everywhere: [z][y][x], x fastest; the spectrum is the
// half spectrum along x (nxh = nx/2 + 1 columns).
// ---------------------------------------------------------------------------

//! Line pairs of the x passes stay inside one xy plane: pair p is plane z = p / pp,
//! pp = ceil(ny / 2), lines y = 2 (p - z pp) and y + 1 -- the second absent (l1 = -1) when ny
//! is odd and y + 1 = ny; a pair past the last plane has l0 = l1 = -1.  The fused xy-plane
//! passes pair the same lines, so both forms compute the same transforms.
struct XPairs
    {
    int ny, pp, npairs;
    FastDiv by_pp;
    __device__ __forceinline__ explicit XPairs(const MeshGeom& g) : ny(g.ny), pp((g.ny + 1) / 2), npairs(((g.ny + 1) / 2) * g.nz), by_pp((g.ny + 1) / 2) {}
    __device__ __forceinline__ void lines(int p, int& l0, int& l1) const
        {
        if (p >= npairs)
            {
            l0 = l1 = -1;
            return;
            }
        const int z = by_pp.quot(p);
        const int y = 2 * (p - z * pp);
        l0 = z * ny + y;
        l1 = (y + 1 < ny) ? l0 + 1 : -1;
        }
    };

//! R2C along x for pairs of real lines packed into one complex FFT; reads and clears the mesh.
template<typename R>
__device__ __forceinline__ void fft_x_r2c_block(int blk,
                                                const MeshGeom& g,
                                                const FFTPlan& plan,
                                                int pairs_per_block,
                                                const C2<R>* __restrict__ tw_g,
                                                typename Acc<R>::glob_t* __restrict__ mesh,
                                                C2<R>* __restrict__ spec,
                                                const float* __restrict__ fx)
    {
    extern __shared__ __align__(16) unsigned char smem[];
    const int nx = g.nx;
    const int ls = nx + 1;
    const double inv_scale = double(fx[1]);
    const FastDiv by_nx(nx), by_nxh(g.nxh);
    C2<R>* tw = reinterpret_cast<C2<R>*>(smem);
    C2<R>* a = tw + nx;
    C2<R>* b = a + pairs_per_block * ls;
    const XPairs xp(g);
    const int pair0 = blk * pairs_per_block;
    for (int i = threadIdx.x; i < nx; i += blockDim.x)
        tw[i] = tw_g[i];
    staged_copy<8, C2<R>>(
        pairs_per_block * nx, threadIdx.x, blockDim.x,
        [&](int idx)
        {
            const int pr = by_nx.quot(idx);
            const int x = idx - pr * nx;
            int l0, l1;
            xp.lines(pair0 + pr, l0, l1);
            R v0 = R(0), v1 = R(0);
            if (l0 >= 0)
                v0 = mesh_load_clear<R>(&mesh[size_t(l0) * nx + x], inv_scale);
            if (l1 >= 0)
                v1 = mesh_load_clear<R>(&mesh[size_t(l1) * nx + x], inv_scale);
            return C2<R> {v0, v1};
        },
        [&](int idx, const C2<R>& v)
        {
            const int pr = by_nx.quot(idx);
            a[pr * ls + idx - pr * nx] = v;
        });
    __syncthreads();
    C2<R>* res = lds_fft<R>(a, b, tw, plan, ls, pairs_per_block, -1, threadIdx.x, blockDim.x);
    const int nxh = g.nxh;
    for (int idx = threadIdx.x; idx < pairs_per_block * nxh; idx += blockDim.x)
        {
        const int pr = by_nxh.quot(idx);
        const int k = idx - pr * nxh;
        int l0, l1;
        xp.lines(pair0 + pr, l0, l1);
        const C2<R> z = res[pr * ls + k];
        const C2<R> zc = res[pr * ls + (k == 0 ? 0 : nx - k)];
        // X0 = (Z[k] + conj(Z[-k])) / 2, X1 = (Z[k] - conj(Z[-k])) / (2 i)
        const C2<R> x0 = {R(0.5) * (z.x + zc.x), R(0.5) * (z.y - zc.y)};
        const C2<R> x1 = {R(0.5) * (z.y + zc.y), R(-0.5) * (z.x - zc.x)};
        if (l0 >= 0)
            spec[size_t(l0) * g.sxp + k] = x0;
        if (l1 >= 0)
            spec[size_t(l1) * g.sxp + k] = x1;
        }
    }

template<typename R>
__global__ void __launch_bounds__(256) k_fft_x_r2c(MeshGeom g,
                                                   FFTPlan plan,
                                                   int pairs_per_block,
                                                   const C2<R>* __restrict__ tw_g,
                                                   typename Acc<R>::glob_t* __restrict__ mesh,
                                                   C2<R>* __restrict__ spec,
                                                   const float* __restrict__ fx)
    {
    fft_x_r2c_block<R>(blockIdx.x, g, plan, pairs_per_block, tw_g, mesh, spec, fx);
    }

//! C2R along x for pairs of half-spectrum lines; phi = scale * real(IFFT).
template<typename R>
__device__ __forceinline__ void fft_x_c2r_block(int blk,
                                                const MeshGeom& g,
                                                const FFTPlan& plan,
                                                int pairs_per_block,
                                                const C2<R>* __restrict__ tw_g,
                                                const C2<R>* __restrict__ spec,
                                                R scale,
                                                R* __restrict__ phi)
    {
    extern __shared__ __align__(16) unsigned char smem[];
    const int nx = g.nx;
    const int nxh = g.nxh;
    const int ls = nx + 1;
    const FastDiv by_nx(nx), by_nxh(nxh);
    C2<R>* tw = reinterpret_cast<C2<R>*>(smem);
    C2<R>* a = tw + nx;
    C2<R>* b = a + pairs_per_block * ls;
    const XPairs xp(g);
    const int pair0 = blk * pairs_per_block;
    for (int i = threadIdx.x; i < nx; i += blockDim.x)
        tw[i] = tw_g[i];
    struct Pair2
        {
        C2<R> x0, x1;
        };
    staged_copy<8, Pair2>(
        pairs_per_block * nxh, threadIdx.x, blockDim.x,
        [&](int idx)
        {
            const int pr = by_nxh.quot(idx);
            const int k = idx - pr * nxh;
            int l0, l1;
            xp.lines(pair0 + pr, l0, l1);
            Pair2 v = {{R(0), R(0)}, {R(0), R(0)}};
            if (l0 >= 0)
                v.x0 = spec[size_t(l0) * g.sxp + k];
            if (l1 >= 0)
                v.x1 = spec[size_t(l1) * g.sxp + k];
            return v;
        },
        [&](int idx, const Pair2& v)
        {
            const int pr = by_nxh.quot(idx);
            const int k = idx - pr * nxh;
            // Z[k] = X0[k] + i X1[k]; Z[n-k] = conj(X0[k]) + i conj(X1[k])
            a[pr * ls + k] = C2<R> {v.x0.x - v.x1.y, v.x0.y + v.x1.x};
            if (k != 0 && k != nx - k)
                a[pr * ls + nx - k] = C2<R> {v.x0.x + v.x1.y, -v.x0.y + v.x1.x};
        });
    __syncthreads();
    C2<R>* res = lds_fft<R>(a, b, tw, plan, ls, pairs_per_block, +1, threadIdx.x, blockDim.x);
    for (int idx = threadIdx.x; idx < pairs_per_block * nx; idx += blockDim.x)
        {
        const int pr = by_nx.quot(idx);
        const int x = idx - pr * nx;
        int l0, l1;
        xp.lines(pair0 + pr, l0, l1);
        const C2<R> z = res[pr * ls + x];
        if (l0 >= 0)
            phi[size_t(l0) * nx + x] = z.x * scale;
        if (l1 >= 0)
            phi[size_t(l1) * nx + x] = z.y * scale;
        }
    }

template<typename R>
__global__ void __launch_bounds__(256) k_fft_x_c2r(MeshGeom g,
                                                   FFTPlan plan,
                                                   int pairs_per_block,
                                                   const C2<R>* __restrict__ tw_g,
                                                   const C2<R>* __restrict__ spec,
                                                   R scale,
                                                   R* __restrict__ phi)
    {
    fft_x_c2r_block<R>(blockIdx.x, g, plan, pairs_per_block, tw_g, spec, scale, phi);
    }

//! Complex FFT along y (sign -1 forward, +1 inverse) on blocks of ncol consecutive x columns.
template<typename R>
__device__ __forceinline__ void fft_y_block(int blk,
                                            const MeshGeom& g,
                                            const FFTPlan& plan,
                                            int ncol,
                                            const C2<R>* __restrict__ tw_g,
                                            C2<R>* __restrict__ spec,
                                            int sign)
    {
    extern __shared__ __align__(16) unsigned char smem[];
    const int ny = g.ny;
    const int nxh = g.nxh;
    const int ls = ny + 1;
    C2<R>* tw = reinterpret_cast<C2<R>*>(smem);
    C2<R>* a = tw + ny;
    C2<R>* b = a + ncol * ls;
    const int nbx = (nxh + ncol - 1) / ncol;
    const int z = blk / nbx;
    const int x0 = (blk - z * nbx) * ncol;
    const int nc = min(ncol, nxh - x0);
    const FastDiv by_ncol(ncol);
    for (int i = threadIdx.x; i < ny; i += blockDim.x)
        tw[i] = tw_g[i];
    staged_copy<8, C2<R>>(
        ny * ncol, threadIdx.x, blockDim.x,
        [&](int idx)
        {
            const int y = by_ncol.quot(idx);
            const int xi = idx - y * ncol;
            return xi < nc ? spec[(size_t(z) * ny + y) * g.sxp + x0 + xi] : C2<R> {R(0), R(0)};
        },
        [&](int idx, const C2<R>& v)
        {
            const int y = by_ncol.quot(idx);
            const int xi = idx - y * ncol;
            if (xi < nc)
                a[xi * ls + y] = v;
        });
    __syncthreads();
    C2<R>* res = lds_fft<R>(a, b, tw, plan, ls, nc, sign, threadIdx.x, blockDim.x);
    for (int idx = threadIdx.x; idx < ny * ncol; idx += blockDim.x)
        {
        const int y = by_ncol.quot(idx);
        const int xi = idx - y * ncol;
        if (xi < nc)
            spec[(size_t(z) * ny + y) * g.sxp + x0 + xi] = res[xi * ls + y];
        }
    }

template<typename R>
__global__ void __launch_bounds__(256) k_fft_y(MeshGeom g, FFTPlan plan, int ncol, const C2<R>* __restrict__ tw_g, C2<R>* __restrict__ spec, int sign)
    {
    fft_y_block<R>(blockIdx.x, g, plan, ncol, tw_g, spec, sign);
    }

//! Forward FFT along z, multiply by the influence function G(k), accumulate the FP64
//! reciprocal energy and virial, inverse FFT along z.  One block per (y, column block);
//! its partial sums go to ev_out[0..6].
template<typename R>
__device__ __forceinline__ void fft_z_conv_block(int blk,
                                                 const MeshGeom& g,
                                                 const FFTPlan& plan,
                                                 int ncol,
                                                 const C2<R>* __restrict__ tw_g,
                                                 C2<R>* __restrict__ spec,
                                                 const double* __restrict__ G,
                                                 const double* __restrict__ Gv,
                                                 int want_ev,
                                                 double* __restrict__ ev_out)
    {
    extern __shared__ __align__(16) unsigned char smem[];
    const int ny = g.ny;
    const int nz = g.nz;
    const int nxh = g.nxh;
    const int ls = nz + 1;
    C2<R>* tw = reinterpret_cast<C2<R>*>(smem);
    C2<R>* a = tw + nz;
    C2<R>* b = a + ncol * ls;
    const int nbx = (nxh + ncol - 1) / ncol;
    const int y = blk / nbx;
    const int x0 = (blk - y * nbx) * ncol;
    const int nc = min(ncol, nxh - x0);
    const FastDiv by_ncol(ncol);
    for (int i = threadIdx.x; i < nz; i += blockDim.x)
        tw[i] = tw_g[i];
    staged_copy<8, C2<R>>(
        nz * ncol, threadIdx.x, blockDim.x,
        [&](int idx)
        {
            const int z = by_ncol.quot(idx);
            const int xi = idx - z * ncol;
            return xi < nc ? spec[(size_t(z) * ny + y) * g.sxp + x0 + xi] : C2<R> {R(0), R(0)};
        },
        [&](int idx, const C2<R>& v)
        {
            const int z = by_ncol.quot(idx);
            const int xi = idx - z * ncol;
            if (xi < nc)
                a[xi * ls + z] = v;
        });
    __syncthreads();
    C2<R>* res = lds_fft<R>(a, b, tw, plan, ls, nc, -1, threadIdx.x, blockDim.x);

    double e = 0, v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
    const double two_pi = 6.283185307179586476925286766559;
    const int ky_i = (y <= ny / 2) ? y : y - ny;
    const double ky0 = two_pi * double(ky_i) / g.Ly;
    struct GG
        {
        double gk, gv;
        };
    staged_copy<8, GG>(
        nz * ncol, threadIdx.x, blockDim.x,
        [&](int idx)
        {
            const int z = by_ncol.quot(idx);
            const int xi = idx - z * ncol;
            GG v = {0.0, 0.0};
            if (xi < nc)
                {
                const size_t m = (size_t(z) * ny + y) * g.sxp + x0 + xi;
                v.gk = G[m];
                if (want_ev)
                    v.gv = Gv[m];
                }
            return v;
        },
        [&](int idx, const GG& gg)
        {
        const int z = by_ncol.quot(idx);
        const int xi = idx - z * ncol;
        if (xi >= nc)
            return;
        const int x = x0 + xi;
        const double gk = gg.gk;
        C2<R> c = res[xi * ls + z];
        if (want_ev)
            {
            const double s2 = double(c.x) * double(c.x) + double(c.y) * double(c.y);
            const double gv = gg.gv;
            const int kz_i = (z <= nz / 2) ? z : z - nz;
            const double kx0 = two_pi * double(x) / g.Lx;
            const double kz0 = two_pi * double(kz_i) / g.Lz;
            double kx, ky, kz;
            mesh_wavevector(g, kx0, ky0, kz0, kx, ky, kz);
            e += s2 * gk;
            v0 += s2 * (gk + gv * kx * kx);
            v1 += s2 * (gv * kx * ky);
            v2 += s2 * (gv * kx * kz);
            v3 += s2 * (gk + gv * ky * ky);
            v4 += s2 * (gv * ky * kz);
            v5 += s2 * (gk + gv * kz * kz);
            if (x != 0 && 2 * x != g.nx)
                {
                // the conjugate mode (-k) outside the half spectrum; a Nyquist index keeps
                // its +N/2 wavenumber (k ordering i <= N/2 -> +i, ESPForceComputeGPU.cu:173-175)
                double kxp, kyp, kzp;
                mesh_wavevector(g, -kx0, (2 * y == ny) ? ky0 : -ky0, (2 * z == nz) ? kz0 : -kz0, kxp, kyp, kzp);
                e += s2 * gk;
                v0 += s2 * (gk + gv * kxp * kxp);
                v1 += s2 * (gv * kxp * kyp);
                v2 += s2 * (gv * kxp * kzp);
                v3 += s2 * (gk + gv * kyp * kyp);
                v4 += s2 * (gv * kyp * kzp);
                v5 += s2 * (gk + gv * kzp * kzp);
                }
            }
        const R gr = R(gk);
        c.x *= gr;
        c.y *= gr;
        res[xi * ls + z] = c;
        });
    __syncthreads();
    C2<R>* other = (res == a) ? b : a;
    C2<R>* res2 = lds_fft<R>(res, other, tw, plan, ls, nc, +1, threadIdx.x, blockDim.x);
    for (int idx = threadIdx.x; idx < nz * ncol; idx += blockDim.x)
        {
        const int z = by_ncol.quot(idx);
        const int xi = idx - z * ncol;
        if (xi < nc)
            spec[(size_t(z) * ny + y) * g.sxp + x0 + xi] = res2[xi * ls + z];
        }
    if (want_ev)
        {
        __shared__ double red[4][7];
        double vals[7] = {e, v0, v1, v2, v3, v4, v5};
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
        for (int c = 0; c < 7; ++c)
            {
            const double s = wave_sum(vals[c]);
            if (lane == 0)
                red[wave][c] = s;
            }
        __syncthreads();
        if (threadIdx.x < 7)
            {
            double s = 0;
            for (int w = 0; w < int(blockDim.x >> 6); ++w)
                s += red[w][threadIdx.x];
            ev_out[threadIdx.x] = s;
            }
        }
    }

template<typename R>
__global__ void __launch_bounds__(256) k_fft_z_conv(MeshGeom g,
                                                    FFTPlan plan,
                                                    int ncol,
                                                    const C2<R>* __restrict__ tw_g,
                                                    C2<R>* __restrict__ spec,
                                                    const double* __restrict__ G,
                                                    const double* __restrict__ Gv,
                                                    int want_ev,
                                                    double* __restrict__ ev_part)
    {
    fft_z_conv_block<R>(blockIdx.x, g, plan, ncol, tw_g, spec, G, Gv, want_ev, ev_part + size_t(blockIdx.x) * 8);
    }

// ---------------------------------------------------------------------------
// Influence function tabulation (FP64): G = F(|k|) / (k^2 |w(kx)|^2 |w(ky)|^2 |w(kz)|^2),
// Gv = (dF/dk / k^3 - 2 F / k^4) / (same window product), both 0 at k = 0 and beyond the
// spherical cutoff |k| > k_cut (k_cut > 0): past the band limit of the split the window
// deconvolution only amplifies round-off of the spread mesh.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_build_influence(MeshGeom g,
                                                         int split_kind,
                                                         double kappa,
                                                         double rc,
                                                         double k_cut,
                                                         double alpha,
                                                         int nq,
                                                         const double* __restrict__ qx,
                                                         const double* __restrict__ qw,
                                                         const double* __restrict__ w2x,
                                                         const double* __restrict__ w2y,
                                                         const double* __restrict__ w2z,
                                                         double* __restrict__ G,
                                                         double* __restrict__ Gv)
    {
    const size_t m = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t total = size_t(g.nz) * g.ny * g.nxh;
    if (m >= total)
        return;
    const int x = int(m % g.nxh);
    const int y = int((m / g.nxh) % g.ny);
    const int z = int(m / (size_t(g.nxh) * g.ny));
    const size_t out = (size_t(z) * g.ny + y) * g.sxp + x;  // padded row layout
    const int ky_i = (y <= g.ny / 2) ? y : y - g.ny;
    const int kz_i = (z <= g.nz / 2) ? z : z - g.nz;
    const double two_pi = 6.283185307179586476925286766559;
    double kx, ky, kz;
    mesh_wavevector(g, two_pi * x / g.Lx, two_pi * ky_i / g.Ly, two_pi * kz_i / g.Lz, kx, ky, kz);
    const double k2 = kx * kx + ky * ky + kz * kz;
    if (k2 == 0.0 || (k_cut > 0.0 && k2 > k_cut * k_cut))
        {
        G[out] = 0.0;
        Gv[out] = 0.0;
        return;
        }
    const double k = sqrt(k2);
    double F = 0.0, dF = 0.0;
    if (split_kind == 1)
        {
        F = exp(-k2 / (4.0 * kappa * kappa));
        dF = -k / (2.0 * kappa * kappa) * F;
        }
    else
        {
        for (int i = 0; i < nq; ++i)
            {
            double s, c;
            sincos(k * rc * qx[i], &s, &c);
            F += qw[i] * c;
            dF -= qw[i] * rc * qx[i] * s;
            }
        }
    if (alpha > 0.0)
        {
        // screening factor exp(-alpha |k|) of the long-range part (ESPForceComputeGPU.cu:199);
        // d/dk of F exp(-alpha k) enters the virial
        const double scr = exp(-alpha * k);
        dF = (dF - alpha * F) * scr;
        F *= scr;
        }
    const double wp = w2x[x] * w2y[y] * w2z[z];
    G[out] = F / (k2 * wp);
    Gv[out] = (dF / (k2 * k) - 2.0 * F / (k2 * k2)) / wp;
    }

//! Final FP64 reduction: out[0] = E_rec + E_self + E_bg, out[1..6] = W_rec + background;
//! qsum[0..1] = (sum q, sum q^2).
__device__ __forceinline__ void finalize_one(int n_ev,
                                             const double* __restrict__ ev_part,
                                             int n_q,
                                             const double* __restrict__ qpart,
                                             double inv_2V,
                                             double self_coeff,
                                             double bg_coeff,
                                             double V,
                                             int want_ev,
                                             double* __restrict__ out,
                                             double* __restrict__ qsum)
    {
    __shared__ double red[16][9];
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (want_ev)
        for (int i = threadIdx.x; i < n_ev; i += blockDim.x)
            for (int c = 0; c < 7; ++c)
                acc[c] += ev_part[size_t(i) * 8 + c];
    for (int i = threadIdx.x; i < n_q; i += blockDim.x)
        {
        acc[7] += qpart[3 * i];
        acc[8] += qpart[3 * i + 1];
        }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < 9; ++c)
        {
        const double s = wave_sum(acc[c]);
        if (lane == 0)
            red[wave][c] = s;
        }
    __syncthreads();
    if (threadIdx.x == 0)
        {
        double t[9];
        for (int c = 0; c < 9; ++c)
            {
            t[c] = 0;
            for (int w = 0; w < int(blockDim.x >> 6); ++w)
                t[c] += red[w][c];
            }
        const double Q = t[7], Q2 = t[8];
        qsum[0] = Q;
        qsum[1] = Q2;
        if (want_ev && out)
            {
            const double Eself = -0.5 * 0.079577471545947673 * self_coeff * Q2;
            const double Ebg = -Q * Q / (2.0 * V) * bg_coeff;
            out[0] = inv_2V * t[0] + Eself + Ebg;
            for (int c = 1; c < 7; ++c)
                out[c] = inv_2V * t[c];
            out[1] += Ebg;
            out[4] += Ebg;
            out[6] += Ebg;
            }
        }
    }

__global__ void __launch_bounds__(256) k_finalize(int n_ev,
                                                  const double* __restrict__ ev_part,
                                                  int n_q,
                                                  const double* __restrict__ qpart,
                                                  double inv_2V,
                                                  double self_coeff,
                                                  double bg_coeff,
                                                  double V,
                                                  int want_ev,
                                                  double* __restrict__ out,
                                                  double* __restrict__ qsum)
    {
    finalize_one(n_ev, ev_part, n_q, qpart, inv_2V, self_coeff, bg_coeff, V, want_ev, out, qsum);
    }

// ---------------------------------------------------------------------------
// Batched stages: one launch covers the work items of every system in a batch.  Each block
// walks the item list with a grid stride; an item names its system and a system-local
// block index, so every item runs exactly the code of the single-system kernel.
// ---------------------------------------------------------------------------
using kernel::BatchItem;
using kernel::BatchList;
using kernel::BatchSys;

//! Tile key of group member i (0xffffffff for an uncharged particle, which is not binned);
//! returns its charge.
__device__ __forceinline__ double bin_key(unsigned int i,
                                          const unsigned int* __restrict__ group,
                                          const double4* __restrict__ postype,
                                          const double* __restrict__ charge,
                                          const MeshGeom& g,
                                          int P,
                                          unsigned int& k)
    {
    const unsigned int p = group ? group[i] : i;
    const double4 r = postype[p];
    const double q = charge[p];
    k = 0xffffffffu;
    if (q != 0.0)
        {
        double t;
        const Frac3 u = box_frac(r.x, r.y, r.z, g);
        const int ox = stencil_origin(u.x, g.nx, P, t);
        const int oy = stencil_origin(u.y, g.ny, P, t);
        const int oz = stencil_origin(u.z, g.nz, P, t);
        k = unsigned((oz / g.T) * g.ty + oy / g.T) * unsigned(g.tx) + unsigned(ox / g.T);
        }
    return q;
    }

//! Batched binning.  A work item is 256 particles of one system; when the system has at most
//! KB_BIN_LDS_TILES tiles (small meshes: batch tiles are 12^3, so a 16^3-64^3 mesh has 8-216
//! tiles) the item counts its particles per tile with LDS atomics and takes one global atomic
//! per occupied tile, instead of one per particle on a handful of hot counters.  Slots within a
//! tile then differ from the single-system order, which changes no result: the mixed-precision
//! mesh is a fixed-point sum and the gather works per particle.
constexpr unsigned int KB_BIN_LDS_TILES = 1024;
__global__ void __launch_bounds__(256) kb_bin(const BatchSys* __restrict__ ds, BatchList l)
    {
    __shared__ double sq[4], sq2[4], sqm[4];
    __shared__ unsigned int lcount[KB_BIN_LDS_TILES];
    for (unsigned int it = blockIdx.x; it < l.n; it += gridDim.x)
        {
        const BatchItem w = l.items[it];
        const BatchSys& d = ds[w.sys];
        const unsigned int i = w.local * kernel::BATCH_CHUNK + threadIdx.x;
        double q = 0.0;
        if (d.n_tiles <= KB_BIN_LDS_TILES)
            {
            for (unsigned int t = threadIdx.x; t < d.n_tiles; t += blockDim.x)
                lcount[t] = 0u;
            __syncthreads();
            unsigned int k = 0xffffffffu, rank = 0;
            if (i < d.n)
                {
                q = bin_key(i, d.group, reinterpret_cast<const double4*>(d.postype), d.charge, d.g, d.P, k);
                if (k != 0xffffffffu)
                    rank = atomicAdd(&lcount[k], 1u);
                }
            __syncthreads();
            // one global atomic per occupied tile: lcount becomes the item's base slot there
            for (unsigned int t = threadIdx.x; t < d.n_tiles; t += blockDim.x)
                {
                const unsigned int c = lcount[t];
                if (c)
                    lcount[t] = atomicAdd(&d.tile_count[t], c);
                }
            __syncthreads();
            if (i < d.n)
                {
                if (k != 0xffffffffu)
                    d.slot[i] = lcount[k] + rank;
                d.key[i] = k;
                }
            }
        else if (i < d.n)
            q = bin_one(i, d.group, reinterpret_cast<const double4*>(d.postype), d.charge, d.g, d.P, d.tile_count, d.key, d.slot);
        const double a = wave_sum(q), b = wave_sum(q * q), c = wave_max(fabs(q));
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        if (lane == 0)
            {
            sq[wave] = a;
            sq2[wave] = b;
            sqm[wave] = c;
            }
        __syncthreads();
        if (threadIdx.x == 0)
            {
            // same summation order as k_bin, so batched and single results agree bit for bit
            double s = 0, s2 = 0, sm = 0;
            for (int k = 0; k < 4; ++k)
                {
                s += sq[k];
                s2 += sq2[k];
                sm = fmax(sm, sqm[k]);
                }
            d.qpart[3 * w.local] = s;
            d.qpart[3 * w.local + 1] = s2;
            d.qpart[3 * w.local + 2] = sm;
            }
        __syncthreads();
        }
    }

__global__ void __launch_bounds__(1024) kb_scan(const BatchSys* __restrict__ ds, BatchList l)
    {
    for (unsigned int it = blockIdx.x; it < l.n; it += gridDim.x)
        {
        const BatchSys& d = ds[l.items[it].sys];
        scan_one(d.n_tiles, d.tile_count, d.tile_start, int(d.n_qblk), d.qpart, d.fxr, d.fx);
        __syncthreads();
        }
    }

template<typename R>
__global__ void __launch_bounds__(256) kb_scatter(const BatchSys* __restrict__ ds, BatchList l)
    {
    for (unsigned int it = blockIdx.x; it < l.n; it += gridDim.x)
        {
        const BatchItem w = l.items[it];
        const BatchSys& d = ds[w.sys];
        const unsigned int i = w.local * kernel::BATCH_CHUNK + threadIdx.x;
        if (i < d.n)
            scatter_one<R>(i, d.group, reinterpret_cast<const double4*>(d.postype), d.charge, d.g, d.P, d.key, d.slot, d.tile_start,
                           reinterpret_cast<SortedRec<R>*>(d.rec), d.sidx);
        }
    }

template<typename R>
__global__ void __launch_bounds__(256) kb_fft_x_r2c(const BatchSys* __restrict__ ds, BatchList l)
    {
    for (unsigned int it = blockIdx.x; it < l.n; it += gridDim.x)
        {
        const BatchItem w = l.items[it];
        const BatchSys& d = ds[w.sys];
        fft_x_r2c_block<R>(int(w.local), d.g, d.plan[0], d.ppb_x, reinterpret_cast<const C2<R>*>(d.tw[0]),
                           reinterpret_cast<typename Acc<R>::glob_t*>(d.mesh), reinterpret_cast<C2<R>*>(d.spec), d.fx);
        __syncthreads();
        }
    }

template<typename R>
__global__ void __launch_bounds__(256) kb_fft_y(const BatchSys* __restrict__ ds, BatchList l, int sign)
    {
    for (unsigned int it = blockIdx.x; it < l.n; it += gridDim.x)
        {
        const BatchItem w = l.items[it];
        const BatchSys& d = ds[w.sys];
        fft_y_block<R>(int(w.local), d.g, d.plan[1], d.ncol_y, reinterpret_cast<const C2<R>*>(d.tw[1]), reinterpret_cast<C2<R>*>(d.spec),
                       sign);
        __syncthreads();
        }
    }

template<typename R>
__global__ void __launch_bounds__(256) kb_fft_z_conv(const BatchSys* __restrict__ ds, BatchList l)
    {
    for (unsigned int it = blockIdx.x; it < l.n; it += gridDim.x)
        {
        const BatchItem w = l.items[it];
        const BatchSys& d = ds[w.sys];
        fft_z_conv_block<R>(int(w.local), d.g, d.plan[2], d.ncol_z, reinterpret_cast<const C2<R>*>(d.tw[2]),
                            reinterpret_cast<C2<R>*>(d.spec), d.G, d.Gv, d.want_ev, d.evpart + size_t(w.local) * 8);
        __syncthreads();
        }
    }

template<typename R>
__global__ void __launch_bounds__(256) kb_fft_x_c2r(const BatchSys* __restrict__ ds, BatchList l)
    {
    for (unsigned int it = blockIdx.x; it < l.n; it += gridDim.x)
        {
        const BatchItem w = l.items[it];
        const BatchSys& d = ds[w.sys];
        fft_x_c2r_block<R>(int(w.local), d.g, d.plan[0], d.ppb_x, reinterpret_cast<const C2<R>*>(d.tw[0]),
                           reinterpret_cast<const C2<R>*>(d.spec), R(1.0 / d.V), reinterpret_cast<R*>(d.phi));
        __syncthreads();
        }
    }

//! Whole xy planes of a small mesh in one workgroup: the x pass (R2C of line pairs, or C2R) and
//! the y pass run back to back in LDS, so the half spectrum makes one round trip to HBM per
//! direction instead of two.  Each 1-D transform is the one the separate passes compute (the
//! same line pairs (XPairs), Stockham stages and twiddles), so the results are bit-identical
//! to fft_x_r2c_block followed by fft_y_block (and fft_y_block(+1) followed by
//! fft_x_c2r_block).  LDS: the two twiddle tables and two buffers of
//! max(ceil(ny/2) (nx + 1), nxh (ny + 1)) points.
template<typename R>
__device__ __forceinline__ void fft_xy_forward_plane(int z,
                                                     const MeshGeom& g,
                                                     const FFTPlan& px,
                                                     const FFTPlan& py,
                                                     const C2<R>* __restrict__ twx_g,
                                                     const C2<R>* __restrict__ twy_g,
                                                     typename Acc<R>::glob_t* __restrict__ mesh,
                                                     C2<R>* __restrict__ spec,
                                                     const float* __restrict__ fx)
    {
    extern __shared__ __align__(16) unsigned char smem[];
    const int nx = g.nx, ny = g.ny, nxh = g.nxh;
    const int np = (ny + 1) / 2;
    const int lsx = nx + 1, lsy = ny + 1;
    const int buf = max(np * lsx, nxh * lsy);
    const double inv_scale = double(fx[1]);
    const FastDiv by_nx(nx), by_nxh(nxh);
    C2<R>* twx = reinterpret_cast<C2<R>*>(smem);
    C2<R>* twy = twx + nx;
    C2<R>* a = twy + ny;
    C2<R>* b = a + buf;
    for (int i = threadIdx.x; i < nx; i += blockDim.x)
        twx[i] = twx_g[i];
    for (int i = threadIdx.x; i < ny; i += blockDim.x)
        twy[i] = twy_g[i];
    typename Acc<R>::glob_t* plane = mesh + size_t(z) * ny * nx;
    staged_copy<8, C2<R>>(
        np * nx, threadIdx.x, blockDim.x,
        [&](int idx)
        {
            const int pr = by_nx.quot(idx);
            const int x = idx - pr * nx;
            const R v0 = mesh_load_clear<R>(&plane[size_t(2 * pr) * nx + x], inv_scale);
            const R v1 = (2 * pr + 1 < ny) ? mesh_load_clear<R>(&plane[size_t(2 * pr + 1) * nx + x], inv_scale) : R(0);
            return C2<R> {v0, v1};
        },
        [&](int idx, const C2<R>& v)
        {
            const int pr = by_nx.quot(idx);
            a[pr * lsx + idx - pr * nx] = v;
        });
    __syncthreads();
    C2<R>* res = lds_fft<R>(a, b, twx, px, lsx, np, -1, threadIdx.x, blockDim.x);
    // unpack the pairs into y lines (one per wavenumber k) of the other buffer
    C2<R>* col = (res == a) ? b : a;
    for (int idx = threadIdx.x; idx < np * nxh; idx += blockDim.x)
        {
        const int pr = by_nxh.quot(idx);
        const int k = idx - pr * nxh;
        const C2<R> zk = res[pr * lsx + k];
        const C2<R> zc = res[pr * lsx + (k == 0 ? 0 : nx - k)];
        col[k * lsy + 2 * pr] = C2<R> {R(0.5) * (zk.x + zc.x), R(0.5) * (zk.y - zc.y)};
        if (2 * pr + 1 < ny)
            col[k * lsy + 2 * pr + 1] = C2<R> {R(0.5) * (zk.y + zc.y), R(-0.5) * (zk.x - zc.x)};
        }
    __syncthreads();
    C2<R>* other = (col == a) ? b : a;
    C2<R>* out = lds_fft<R>(col, other, twy, py, lsy, nxh, -1, threadIdx.x, blockDim.x);
    C2<R>* sp = spec + size_t(z) * ny * g.sxp;
    for (int idx = threadIdx.x; idx < ny * nxh; idx += blockDim.x)
        {
        const int y = by_nxh.quot(idx);
        const int k = idx - y * nxh;
        sp[size_t(y) * g.sxp + k] = out[k * lsy + y];
        }
    }

template<typename R>
__device__ __forceinline__ void fft_xy_inverse_plane(int z,
                                                     const MeshGeom& g,
                                                     const FFTPlan& px,
                                                     const FFTPlan& py,
                                                     const C2<R>* __restrict__ twx_g,
                                                     const C2<R>* __restrict__ twy_g,
                                                     const C2<R>* __restrict__ spec,
                                                     R scale,
                                                     R* __restrict__ phi)
    {
    extern __shared__ __align__(16) unsigned char smem[];
    const int nx = g.nx, ny = g.ny, nxh = g.nxh;
    const int np = (ny + 1) / 2;
    const int lsx = nx + 1, lsy = ny + 1;
    const int buf = max(np * lsx, nxh * lsy);
    const FastDiv by_nx(nx), by_nxh(nxh);
    C2<R>* twx = reinterpret_cast<C2<R>*>(smem);
    C2<R>* twy = twx + nx;
    C2<R>* a = twy + ny;
    C2<R>* b = a + buf;
    for (int i = threadIdx.x; i < nx; i += blockDim.x)
        twx[i] = twx_g[i];
    for (int i = threadIdx.x; i < ny; i += blockDim.x)
        twy[i] = twy_g[i];
    const C2<R>* sp = spec + size_t(z) * ny * g.sxp;
    staged_copy<8, C2<R>>(
        ny * nxh, threadIdx.x, blockDim.x,
        [&](int idx)
        {
            const int y = by_nxh.quot(idx);
            return sp[size_t(y) * g.sxp + idx - y * nxh];
        },
        [&](int idx, const C2<R>& v)
        {
            const int y = by_nxh.quot(idx);
            a[(idx - y * nxh) * lsy + y] = v;
        });
    __syncthreads();
    C2<R>* res = lds_fft<R>(a, b, twy, py, lsy, nxh, +1, threadIdx.x, blockDim.x);
    // Z[k] = X0[k] + i X1[k]; Z[n-k] = conj(X0[k]) + i conj(X1[k]) for the line pair (2 pr, 2 pr + 1)
    C2<R>* lines = (res == a) ? b : a;
    for (int idx = threadIdx.x; idx < np * nxh; idx += blockDim.x)
        {
        const int pr = by_nxh.quot(idx);
        const int k = idx - pr * nxh;
        const C2<R> x0 = res[k * lsy + 2 * pr];
        const C2<R> x1 = (2 * pr + 1 < ny) ? res[k * lsy + 2 * pr + 1] : C2<R> {R(0), R(0)};
        lines[pr * lsx + k] = C2<R> {x0.x - x1.y, x0.y + x1.x};
        if (k != 0 && k != nx - k)
            lines[pr * lsx + nx - k] = C2<R> {x0.x + x1.y, -x0.y + x1.x};
        }
    __syncthreads();
    C2<R>* other = (lines == a) ? b : a;
    C2<R>* out = lds_fft<R>(lines, other, twx, px, lsx, np, +1, threadIdx.x, blockDim.x);
    R* ph = phi + size_t(z) * ny * nx;
    for (int idx = threadIdx.x; idx < np * nx; idx += blockDim.x)
        {
        const int pr = by_nx.quot(idx);
        const int x = idx - pr * nx;
        const C2<R> v = out[pr * lsx + x];
        ph[size_t(2 * pr) * nx + x] = v.x * scale;
        if (2 * pr + 1 < ny)
            ph[size_t(2 * pr + 1) * nx + x] = v.y * scale;
        }
    }

template<typename R>
__global__ void __launch_bounds__(256) kb_fft_xy_forward(const BatchSys* __restrict__ ds, BatchList l)
    {
    for (unsigned int it = blockIdx.x; it < l.n; it += gridDim.x)
        {
        const BatchItem w = l.items[it];
        const BatchSys& d = ds[w.sys];
        fft_xy_forward_plane<R>(int(w.local), d.g, d.plan[0], d.plan[1], reinterpret_cast<const C2<R>*>(d.tw[0]),
                                reinterpret_cast<const C2<R>*>(d.tw[1]), reinterpret_cast<typename Acc<R>::glob_t*>(d.mesh),
                                reinterpret_cast<C2<R>*>(d.spec), d.fx);
        __syncthreads();
        }
    }

template<typename R>
__global__ void __launch_bounds__(256) kb_fft_xy_inverse(const BatchSys* __restrict__ ds, BatchList l)
    {
    for (unsigned int it = blockIdx.x; it < l.n; it += gridDim.x)
        {
        const BatchItem w = l.items[it];
        const BatchSys& d = ds[w.sys];
        fft_xy_inverse_plane<R>(int(w.local), d.g, d.plan[0], d.plan[1], reinterpret_cast<const C2<R>*>(d.tw[0]),
                                reinterpret_cast<const C2<R>*>(d.tw[1]), reinterpret_cast<const C2<R>*>(d.spec), R(1.0 / d.V),
                                reinterpret_cast<R*>(d.phi));
        __syncthreads();
        }
    }

__global__ void __launch_bounds__(256) kb_finalize(const BatchSys* __restrict__ ds, BatchList l)
    {
    for (unsigned int it = blockIdx.x; it < l.n; it += gridDim.x)
        {
        const BatchSys& d = ds[l.items[it].sys];
        finalize_one(int(d.n_zblk), d.evpart, int(d.n_qblk), d.qpart, d.inv_2V, d.self_coeff, d.bg_coeff, d.V, d.want_ev, d.ev_out,
                     d.qsum);
        __syncthreads();
        }
    }

    } // namespace dev

// ===========================================================================
// Host launchers (the kernel-launcher layer; cf. ESPForceComputeGPU.cuh:49-144)
// ===========================================================================
namespace kernel
{
using namespace esp::dev;

hipError_t bin_particles(unsigned int n,
                         const unsigned int* group,
                         const double* postype,
                         const double* charge,
                         const MeshGeom& g,
                         int P,
                         unsigned int* tile_count,
                         unsigned int* key,
                         unsigned int* slot,
                         double* qpart,
                         hipStream_t s)
    {
    const unsigned int nb = (n + 255) / 256;
    if (nb == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_bin, dim3(nb), dim3(256), 0, s, n, group, reinterpret_cast<const double4*>(postype), charge, g, P,
                       tile_count, key, slot, qpart);
    return hipGetLastError();
    }

hipError_t scan_tiles(unsigned int n_tiles, unsigned int* count, unsigned int* start, int n_q, const double* qpart, FixedPointRule rule,
                      float* fx, hipStream_t s)
    {
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, n_tiles, count, start, n_q, qpart, rule, fx);
    return hipGetLastError();
    }

template<typename R>
hipError_t scatter_particles(unsigned int n,
                             const unsigned int* group,
                             const double* postype,
                             const double* charge,
                             const MeshGeom& g,
                             int P,
                             const unsigned int* key,
                             const unsigned int* slot,
                             const unsigned int* start,
                             void* rec,
                             SortedIdx* sidx,
                             hipStream_t s)
    {
    const unsigned int nb = (n + 255) / 256;
    if (nb == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_scatter<R>, dim3(nb), dim3(256), 0, s, n, group, reinterpret_cast<const double4*>(postype), charge, g,
                       P, key, slot, start, reinterpret_cast<SortedRec<R>*>(rec), sidx);
    return hipGetLastError();
    }
template hipError_t scatter_particles<double>(unsigned int, const unsigned int*, const double*, const double*, const MeshGeom&, int,
                                              const unsigned int*, const unsigned int*, const unsigned int*, void*, SortedIdx*,
                                              hipStream_t);
template hipError_t scatter_particles<float>(unsigned int, const unsigned int*, const double*, const double*, const MeshGeom&, int,
                                             const unsigned int*, const unsigned int*, const unsigned int*, void*, SortedIdx*,
                                             hipStream_t);

int fft_x_pairs_per_block(int nx, bool fp64)
    {
    const size_t el = fp64 ? 16 : 8;
    int pairs = 8;
    while (pairs > 1 && (size_t(nx) + 2 * size_t(pairs) * (nx + 1)) * el > 64 * 1024)
        pairs >>= 1;
    return pairs;
    }

//! Columns per workgroup of the y and z passes (lines of length n over the nxh columns of the
//! half spectrum): at most 16 within 64 KiB of LDS, then evened out over the blocks a row of
//! nxh columns needs (nxh = 33 takes three blocks of 11, not 16 + 16 + 1), so no block runs its
//! copy loops over mostly absent columns.
int fft_cols_per_block(int n, int nxh, bool fp64)
    {
    const size_t el = fp64 ? 16 : 8;
    int cols = 16;
    while (cols > 1 && (size_t(n) + 2 * size_t(cols) * (n + 1)) * el > 64 * 1024)
        cols >>= 1;
    const int blocks = (nxh + cols - 1) / cols;
    return (nxh + blocks - 1) / blocks;
    }

template<typename R>
hipError_t fft_forward_x(const MeshGeom& g, const FFTPlan& px, const void* tw, void* mesh, void* spec, const float* fx, hipStream_t s)
    {
    const int ppb = fft_x_pairs_per_block(g.nx, sizeof(R) == 8);
    const int npairs = ((g.ny + 1) / 2) * g.nz;
    const int nb = (npairs + ppb - 1) / ppb;
    const size_t lds = (size_t(g.nx) + 2 * size_t(ppb) * (g.nx + 1)) * sizeof(C2<R>);
    hipLaunchKernelGGL(k_fft_x_r2c<R>, dim3(nb), dim3(256), lds, s, g, px, ppb, reinterpret_cast<const C2<R>*>(tw),
                       reinterpret_cast<typename Acc<R>::glob_t*>(mesh), reinterpret_cast<C2<R>*>(spec), fx);
    return hipGetLastError();
    }

template<typename R>
hipError_t fft_inverse_x(const MeshGeom& g, const FFTPlan& px, const void* tw, const void* spec, double scale, void* phi, hipStream_t s)
    {
    const int ppb = fft_x_pairs_per_block(g.nx, sizeof(R) == 8);
    const int npairs = ((g.ny + 1) / 2) * g.nz;
    const int nb = (npairs + ppb - 1) / ppb;
    const size_t lds = (size_t(g.nx) + 2 * size_t(ppb) * (g.nx + 1)) * sizeof(C2<R>);
    hipLaunchKernelGGL(k_fft_x_c2r<R>, dim3(nb), dim3(256), lds, s, g, px, ppb, reinterpret_cast<const C2<R>*>(tw),
                       reinterpret_cast<const C2<R>*>(spec), R(scale), reinterpret_cast<R*>(phi));
    return hipGetLastError();
    }

template<typename R>
hipError_t fft_y(const MeshGeom& g, const FFTPlan& py, const void* tw, void* spec, int sign, hipStream_t s)
    {
    const int ncol = fft_cols_per_block(g.ny, g.nxh, sizeof(R) == 8);
    const int nbx = (g.nxh + ncol - 1) / ncol;
    const size_t lds = (size_t(g.ny) + 2 * size_t(ncol) * (g.ny + 1)) * sizeof(C2<R>);
    hipLaunchKernelGGL(k_fft_y<R>, dim3(nbx * g.nz), dim3(256), lds, s, g, py, ncol, reinterpret_cast<const C2<R>*>(tw),
                       reinterpret_cast<C2<R>*>(spec), sign);
    return hipGetLastError();
    }

unsigned int fft_z_blocks(const MeshGeom& g, bool fp64)
    {
    const int ncol = fft_cols_per_block(g.nz, g.nxh, fp64);
    const int nbx = (g.nxh + ncol - 1) / ncol;
    return unsigned(nbx * g.ny);
    }

template<typename R>
hipError_t fft_z_convolve(const MeshGeom& g, const FFTPlan& pz, const void* tw, void* spec, const double* G, const double* Gv,
                          int want_ev, double* ev_part, hipStream_t s)
    {
    const int ncol = fft_cols_per_block(g.nz, g.nxh, sizeof(R) == 8);
    const int nbx = (g.nxh + ncol - 1) / ncol;
    const size_t lds = (size_t(g.nz) + 2 * size_t(ncol) * (g.nz + 1)) * sizeof(C2<R>);
    hipLaunchKernelGGL(k_fft_z_conv<R>, dim3(nbx * g.ny), dim3(256), lds, s, g, pz, ncol, reinterpret_cast<const C2<R>*>(tw),
                       reinterpret_cast<C2<R>*>(spec), G, Gv, want_ev, ev_part);
    return hipGetLastError();
    }

#define ESP_INST_FFT(R)                                                                                                          \
    template hipError_t fft_forward_x<R>(const MeshGeom&, const FFTPlan&, const void*, void*, void*, const float*, hipStream_t);                \
    template hipError_t fft_inverse_x<R>(const MeshGeom&, const FFTPlan&, const void*, const void*, double, void*, hipStream_t);  \
    template hipError_t fft_y<R>(const MeshGeom&, const FFTPlan&, const void*, void*, int, hipStream_t);                        \
    template hipError_t fft_z_convolve<R>(const MeshGeom&, const FFTPlan&, const void*, void*, const double*, const double*, int, \
                                          double*, hipStream_t);
ESP_INST_FFT(double)
ESP_INST_FFT(float)
#undef ESP_INST_FFT

hipError_t build_influence(const MeshGeom& g, int split_kind, double kappa, double rc, double k_cut, double alpha, int nq,
                           const double* qx, const double* qw, const double* w2x, const double* w2y, const double* w2z, double* G,
                           double* Gv, hipStream_t s)
    {
    const size_t total = size_t(g.nz) * g.ny * g.nxh;
    const unsigned int nb = unsigned((total + 255) / 256);
    hipLaunchKernelGGL(k_build_influence, dim3(nb), dim3(256), 0, s, g, split_kind, kappa, rc, k_cut, alpha, nq, qx, qw, w2x, w2y, w2z,
                       G, Gv);
    return hipGetLastError();
    }

hipError_t finalize(int n_ev, const double* ev_part, int n_q, const double* qpart, double inv_2V, double self_coeff, double bg_coeff,
                    double V, int want_ev, double* out, double* qsum, hipStream_t s)
    {
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, n_ev, ev_part, n_q, qpart, inv_2V, self_coeff, bg_coeff, V, want_ev, out,
                       qsum);
    return hipGetLastError();
    }

// ---- batched launchers ---------------------------------------------------------------
hipError_t batch_bin(const BatchSys* d, BatchList l, hipStream_t s)
    {
    if (l.n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(kb_bin, dim3(l.grid), dim3(256), 0, s, d, l);
    return hipGetLastError();
    }

hipError_t batch_scan(const BatchSys* d, BatchList l, hipStream_t s)
    {
    if (l.n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(kb_scan, dim3(l.grid), dim3(1024), 0, s, d, l);
    return hipGetLastError();
    }

template<typename R> hipError_t batch_scatter(const BatchSys* d, BatchList l, hipStream_t s)
    {
    if (l.n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(kb_scatter<R>, dim3(l.grid), dim3(256), 0, s, d, l);
    return hipGetLastError();
    }

template<typename R> hipError_t batch_fft_x_r2c(const BatchSys* d, BatchList l, size_t lds, hipStream_t s)
    {
    if (l.n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(kb_fft_x_r2c<R>, dim3(l.grid), dim3(256), lds, s, d, l);
    return hipGetLastError();
    }

template<typename R> hipError_t batch_fft_y(const BatchSys* d, BatchList l, int sign, size_t lds, hipStream_t s)
    {
    if (l.n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(kb_fft_y<R>, dim3(l.grid), dim3(256), lds, s, d, l, sign);
    return hipGetLastError();
    }

template<typename R> hipError_t batch_fft_z_conv(const BatchSys* d, BatchList l, size_t lds, hipStream_t s)
    {
    if (l.n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(kb_fft_z_conv<R>, dim3(l.grid), dim3(256), lds, s, d, l);
    return hipGetLastError();
    }

size_t fft_xy_lds(const MeshGeom& g, bool fp64)
    {
    const size_t buf = std::max(size_t((g.ny + 1) / 2) * size_t(g.nx + 1), size_t(g.nxh) * size_t(g.ny + 1));
    return (size_t(g.nx) + size_t(g.ny) + 2 * buf) * (fp64 ? 16 : 8);
    }

bool fft_xy_fused(const MeshGeom& g, bool fp64)
    {
    return fft_xy_lds(g, fp64) <= 64 * 1024;
    }

template<typename R> hipError_t batch_fft_xy_forward(const BatchSys* d, BatchList l, size_t lds, hipStream_t s)
    {
    if (l.n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(kb_fft_xy_forward<R>, dim3(l.grid), dim3(256), lds, s, d, l);
    return hipGetLastError();
    }

template<typename R> hipError_t batch_fft_xy_inverse(const BatchSys* d, BatchList l, size_t lds, hipStream_t s)
    {
    if (l.n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(kb_fft_xy_inverse<R>, dim3(l.grid), dim3(256), lds, s, d, l);
    return hipGetLastError();
    }

template<typename R> hipError_t batch_fft_x_c2r(const BatchSys* d, BatchList l, size_t lds, hipStream_t s)
    {
    if (l.n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(kb_fft_x_c2r<R>, dim3(l.grid), dim3(256), lds, s, d, l);
    return hipGetLastError();
    }

hipError_t batch_finalize(const BatchSys* d, BatchList l, hipStream_t s)
    {
    if (l.n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(kb_finalize, dim3(l.grid), dim3(256), 0, s, d, l);
    return hipGetLastError();
    }

#define ESP_INST_BATCH(R)                                                                     \
    template hipError_t batch_scatter<R>(const BatchSys*, BatchList, hipStream_t);            \
    template hipError_t batch_fft_x_r2c<R>(const BatchSys*, BatchList, size_t, hipStream_t);  \
    template hipError_t batch_fft_y<R>(const BatchSys*, BatchList, int, size_t, hipStream_t); \
    template hipError_t batch_fft_z_conv<R>(const BatchSys*, BatchList, size_t, hipStream_t); \
    template hipError_t batch_fft_x_c2r<R>(const BatchSys*, BatchList, size_t, hipStream_t);  \
    template hipError_t batch_fft_xy_forward<R>(const BatchSys*, BatchList, size_t, hipStream_t); \
    template hipError_t batch_fft_xy_inverse<R>(const BatchSys*, BatchList, size_t, hipStream_t);
ESP_INST_BATCH(double)
ESP_INST_BATCH(float)
#undef ESP_INST_BATCH

    } // namespace kernel
    } // namespace esp
