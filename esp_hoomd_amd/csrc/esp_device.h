// Device-side building blocks shared by the ESP kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "esp_types.h"

namespace esp
{
namespace dev
{

constexpr int WAVE = 64;

//! Unwrapped fractional coordinates u = B^-1 r + 1/2 of a position r in a box centred on the
//! origin, B = [a1 a2 a3] with a1 = (Lx, 0, 0), a2 = (xy Ly, Ly, 0), a3 = (xz Lz, yz Lz, Lz)
//! (HOOMD's BoxDim::makeFraction).  An orthorhombic box takes the plain x / L + 1/2.
struct Frac3
    {
    double x, y, z;
    };
//! The tilted forms are written with explicit fused multiply-adds, so that every kernel that
//! inlines them (single-system and batched) rounds them the same way.
__device__ __forceinline__ Frac3 box_frac(double x, double y, double z, const MeshGeom& g)
    {
    if (!g.tilted)
        return {x / g.Lx + 0.5, y / g.Ly + 0.5, z / g.Lz + 0.5};
    const double cxz = fma(-g.xy, g.yz, g.xz);  // xz - xy yz
    return {fma(-cxz, z, fma(-g.xy, y, x)) / g.Lx + 0.5, fma(-g.yz, z, y) / g.Ly + 0.5, z / g.Lz + 0.5};
    }

//! Wavevector k = 2 pi B^-T m of the mesh mode m, given the per-axis wavenumbers
//! k0 = 2 pi m_a / L_a: k_x = k0_x, k_y = k0_y - xy k0_x, k_z = k0_z - yz k0_y + (xy yz - xz) k0_x.
__device__ __forceinline__ void mesh_wavevector(const MeshGeom& g, double k0x, double k0y, double k0z, double& kx, double& ky,
                                                double& kz)
    {
    kx = k0x;
    ky = k0y;
    kz = k0z;
    if (g.tilted)
        {
        ky = fma(-g.xy, k0x, k0y);
        kz = fma(fma(g.xy, g.yz, -g.xz), k0x, fma(-g.yz, k0y, k0z));
        }
    }

//! Cartesian gradient from the gradients along the mesh axes scaled by 1 / h_a
//! (d/dr = B^-T d/du): g_y -= xy g_x, g_z += (xy yz - xz) g_x - yz g_y.
__device__ __forceinline__ void mesh_gradient_cartesian(const MeshGeom& g, double& gx, double& gy, double& gz)
    {
    if (g.tilted)
        {
        gz = fma(fma(g.xy, g.yz, -g.xz), gx, fma(-g.yz, gy, gz));
        gy = fma(-g.xy, gx, gy);
        }
    }

template<typename T> __device__ __forceinline__ T wave_sum(T v)
    {
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1)
        v += __shfl_xor(v, off, WAVE);
    return v;
    }

template<typename T> __device__ __forceinline__ T wave_max(T v)
    {
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1)
        {
        const T o = __shfl_xor(v, off, WAVE);
        v = o > v ? o : v;
        }
    return v;
    }

//! Sum over aligned groups of G lanes (G a power of two <= 64).
template<int G, typename T> __device__ __forceinline__ T group_sum(T v)
    {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1)
        v += __shfl_xor(v, off, WAVE);
    return v;
    }

//! Horner in z over D+1 coefficients c[0..D] (c[j] multiplies z^j), value and d/dz.
template<typename R, int D> __device__ __forceinline__ R horner(const R* c, R z)
    {
    R v = c[D];
#pragma unroll
    for (int j = D - 1; j >= 0; --j)
        v = v * z + c[j];
    return v;
    }

template<typename R, int D> __device__ __forceinline__ R horner_d(const R* c, R z, R& dv)
    {
    R v = c[D];
    R d = R(0);
#pragma unroll
    for (int j = D - 1; j >= 0; --j)
        {
        d = d * z + v;
        v = v * z + c[j];
        }
    dv = d;
    return v;
    }

// ---------------------------------------------------------------------------
// Complex helpers on (x, y) pairs
// ---------------------------------------------------------------------------
template<typename R> struct C2
    {
    R x, y;
    };

//! a b + c with one rounding.  The FFT arithmetic is written with these explicitly, so the
//! compiler has no product-sum left to contract: a transform rounds the same way in every
//! kernel it is inlined into, whatever the kernel knows of its sign or sizes (the batched,
//! fused and single-system passes give bit-identical spectra).
template<typename R> __device__ __forceinline__ R fmar(R a, R b, R c)
    {
    if constexpr (sizeof(R) == 4)
        return __fmaf_rn(a, b, c);
    else
        return __fma_rn(a, b, c);
    }

//! a b and a conj(b)
template<typename R> __device__ __forceinline__ C2<R> cmul(C2<R> a, C2<R> b)
    {
    return {fmar(a.x, b.x, -(a.y * b.y)), fmar(a.x, b.y, a.y * b.x)};
    }
template<typename R> __device__ __forceinline__ C2<R> cmul_conj(C2<R> a, C2<R> b)
    {
    return {fmar(a.x, b.x, a.y * b.y), fmar(a.y, b.x, -(a.x * b.y))};
    }
template<typename R> __device__ __forceinline__ C2<R> cadd(C2<R> a, C2<R> b)
    {
    return {a.x + b.x, a.y + b.y};
    }
template<typename R> __device__ __forceinline__ C2<R> csub(C2<R> a, C2<R> b)
    {
    return {a.x - b.x, a.y - b.y};
    }
//! multiply by -i (forward, sign = -1) or +i (inverse, sign = +1)
template<typename R> __device__ __forceinline__ C2<R> cmul_si(C2<R> a, int sign)
    {
    return sign < 0 ? C2<R> {a.y, -a.x} : C2<R> {-a.y, a.x};
    }

// ---------------------------------------------------------------------------
// Radix-R butterflies, in place on v[0..R), sign -1 forward / +1 inverse.
// ---------------------------------------------------------------------------
template<typename R> __device__ __forceinline__ void bfly2(C2<R>* v)
    {
    const C2<R> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
    }

template<typename R> __device__ __forceinline__ void bfly4(C2<R>* v, int sign)
    {
    const C2<R> a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    const C2<R> b0 = cadd(v[1], v[3]), b1 = cmul_si(csub(v[1], v[3]), sign);
    v[0] = cadd(a0, b0);
    v[2] = csub(a0, b0);
    v[1] = cadd(a1, b1);
    v[3] = csub(a1, b1);
    }

template<typename R> __device__ __forceinline__ void bfly3(C2<R>* v, int sign)
    {
    // exp(sign * 2 pi i / 3) = -1/2 + sign * i sqrt(3)/2
    const R c1 = R(-0.5);
    const R s1 = R(sign) * R(0.86602540378443864676);
    const C2<R> t = cadd(v[1], v[2]);
    const C2<R> d = csub(v[1], v[2]);
    const C2<R> m = {fmar(c1, t.x, v[0].x), fmar(c1, t.y, v[0].y)};
    // i * s1 * d
    const C2<R> j = {-s1 * d.y, s1 * d.x};
    v[0] = cadd(v[0], t);
    v[1] = cadd(m, j);
    v[2] = csub(m, j);
    }

template<typename R> __device__ __forceinline__ void bfly5(C2<R>* v, int sign)
    {
    const R c1 = R(0.30901699437494742410);   // cos(2pi/5)
    const R c2 = R(-0.80901699437494742410);  // cos(4pi/5)
    const R s1 = R(sign) * R(0.95105651629515357212);
    const R s2 = R(sign) * R(0.58778525229247312917);
    const C2<R> t1 = cadd(v[1], v[4]), d1 = csub(v[1], v[4]);
    const C2<R> t2 = cadd(v[2], v[3]), d2 = csub(v[2], v[3]);
    const C2<R> a1 = {fmar(c2, t2.x, fmar(c1, t1.x, v[0].x)), fmar(c2, t2.y, fmar(c1, t1.y, v[0].y))};
    const C2<R> a2 = {fmar(c1, t2.x, fmar(c2, t1.x, v[0].x)), fmar(c1, t2.y, fmar(c2, t1.y, v[0].y))};
    // i * (s1 d1 + s2 d2), i * (s2 d1 - s1 d2)
    const C2<R> b1 = {-fmar(s1, d1.y, s2 * d2.y), fmar(s1, d1.x, s2 * d2.x)};
    const C2<R> b2 = {-fmar(s2, d1.y, -(s1 * d2.y)), fmar(s2, d1.x, -(s1 * d2.x))};
    v[0] = {v[0].x + t1.x + t2.x, v[0].y + t1.y + t2.y};
    v[1] = cadd(a1, b1);
    v[4] = csub(a1, b1);
    v[2] = cadd(a2, b2);
    v[3] = csub(a2, b2);
    }

//! Quotient by a divisor fixed for a loop, as one v_mul_hi_u32 instead of the ~20 instructions
//! of a general 32-bit division.  With m = ceil(2^32 / d) = (2^32 + e) / d, 0 <= e < d, and
//! n = q d + r: n m / 2^32 = q + (r + n e / 2^32) / d, whose floor is q while n e < 2^32, which
//! n d <= 2^32 guarantees.  The FFT passes divide indices below 2^20 by divisors of at most
//! 1024 (mesh edges are at most 1024).
//! d = 1 (m would be 2^32) returns n.
struct FastDiv
    {
    unsigned int d, m;
    __device__ __forceinline__ explicit FastDiv(int div) : d(unsigned(div)), m(div > 1 ? 0xffffffffu / unsigned(div) + 1u : 0u) {}
    __device__ __forceinline__ int quot(int n) const
        {
        return d == 1u ? n : int(__umulhi(unsigned(n), m));
        }
    };

//! One Stockham stage over nlines lines of length N stored with line stride ls in LDS.
//! tw[m] = exp(-2 pi i m / N); the inverse uses its conjugate.  The twiddle index
//! k r (N / (Ns RAD)) <= (Ns - 1)(RAD - 1) N / (Ns RAD) stays below N, so it needs no wrap.
template<typename R, int RAD>
__device__ __forceinline__ void stockham_stage(const C2<R>* __restrict__ in,
                                               C2<R>* __restrict__ out,
                                               const C2<R>* __restrict__ tw,
                                               int N,
                                               int ls,
                                               int nlines,
                                               int Ns,
                                               int sign,
                                               int tid,
                                               int nthreads)
    {
    const int nb = N / RAD;
    const int twstep = N / (Ns * RAD);
    const FastDiv by_nb(nb), by_ns(Ns);
    for (int idx = tid; idx < nlines * nb; idx += nthreads)
        {
        const int line = by_nb.quot(idx);
        const int j = idx - line * nb;
        const C2<R>* src = in + line * ls;
        C2<R>* dst = out + line * ls;
        C2<R> v[RAD];
        const int jq = by_ns.quot(j);
        const int k = j - jq * Ns;
#pragma unroll
        for (int r = 0; r < RAD; ++r)
            v[r] = src[j + r * nb];
#pragma unroll
        for (int r = 1; r < RAD; ++r)
            {
            const C2<R> w = tw[k * r * twstep];
            v[r] = sign > 0 ? cmul_conj(v[r], w) : cmul(v[r], w);
            }
        if constexpr (RAD == 2)
            bfly2(v);
        else if constexpr (RAD == 3)
            bfly3(v, sign);
        else if constexpr (RAD == 4)
            bfly4(v, sign);
        else
            bfly5(v, sign);
        const int base = jq * Ns * RAD + k;
#pragma unroll
        for (int r = 0; r < RAD; ++r)
            dst[base + r * Ns] = v[r];
        }
    }

//! for (i = tid; i < n; i += nt) put(i, get(i)), with the get()s of NB consecutive iterations
//! issued before their put()s, so NB global loads are in flight per thread.  (When put() stores
//! through a pointer the compiler cannot prove distinct from get()'s source, a plain loop waits
//! for every load before its store.)  Iterations run in the same order as the plain loop.
template<int NB, typename T, typename Get, typename Put>
__device__ __forceinline__ void staged_copy(int n, int tid, int nt, Get get, Put put)
    {
    for (int i0 = tid; i0 < n; i0 += NB * nt)
        {
        T v[NB];
#pragma unroll
        for (int e = 0; e < NB; ++e)
            if (i0 + e * nt < n)
                v[e] = get(i0 + e * nt);
#pragma unroll
        for (int e = 0; e < NB; ++e)
            if (i0 + e * nt < n)
                put(i0 + e * nt, v[e]);
        }
    }

//! Full 1-D FFT of nlines lines in LDS (ping-pong between a and b).  Returns the buffer holding
//! the result.  radices[0..nstages) multiply to N.
template<typename R>
__device__ C2<R>* lds_fft(C2<R>* a,
                          C2<R>* b,
                          const C2<R>* tw,
                          const FFTPlan& plan,
                          int ls,
                          int nlines,
                          int sign,
                          int tid,
                          int nthreads)
    {
    int Ns = 1;
    C2<R>* src = a;
    C2<R>* dst = b;
    for (int s = 0; s < plan.nstages; ++s)
        {
        const int rad = plan.radix[s];
        if (rad == 4)
            stockham_stage<R, 4>(src, dst, tw, plan.n, ls, nlines, Ns, sign, tid, nthreads);
        else if (rad == 2)
            stockham_stage<R, 2>(src, dst, tw, plan.n, ls, nlines, Ns, sign, tid, nthreads);
        else if (rad == 3)
            stockham_stage<R, 3>(src, dst, tw, plan.n, ls, nlines, Ns, sign, tid, nthreads);
        else
            stockham_stage<R, 5>(src, dst, tw, plan.n, ls, nlines, Ns, sign, tid, nthreads);
        __syncthreads();
        Ns *= rad;
        C2<R>* t = src;
        src = dst;
        dst = t;
        }
    return src;
    }

// ---------------------------------------------------------------------------
// Mesh accumulation
// ---------------------------------------------------------------------------
//! Accumulator types.  Both precisions accumulate the mesh in fixed point, which is exact and
//! independent of the order in which particles arrive: the mixed mode rounds FP32 taps into
//! int32 LDS cells, the FP64 mode rounds FP64 taps into int64 cells, and both meshes in HBM are
//! int64.  The step's scale (the tile scan's fx[0]) keeps every tap and every cell sum inside
//! its integer type, so the spread mesh -- and everything computed from it -- is bit-identical
//! from run to run and for any particle order or tiling of the same positions.
template<typename R> struct Acc;
template<> struct Acc<double>
    {
    using lds_t = long long;
    using glob_t = unsigned long long;
    };
template<> struct Acc<float>
    {
    using lds_t = int;
    using glob_t = unsigned long long;
    };

//! Nearest integer (ties up) in one instruction, v_cvt_rpi_i32_f32 = floor(v + 0.5); the
//! accumulated values stay far inside the int32 range.
__device__ __forceinline__ int round_to_int(float v)
    {
    int r;
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(r) : "v"(v));
    return r;
    }
//! 1.5 * 2^52 and its bit pattern: for |v| < 2^51, v + FIX_MAGIC lies in [2^52, 2^53), where
//! the spacing of doubles is 1, so the sum's bits are FIX_MAGIC_BITS + rint(v) (round to
//! nearest even) -- an int64 conversion in one FP64 add (or fma, which rounds the exact
//! product) and one integer subtract.
constexpr double FIX_MAGIC = 6755399441055744.0;
constexpr long long FIX_MAGIC_BITS = 0x4338000000000000LL;
__device__ __forceinline__ long long round_to_ll(double v)
    {
    return __double_as_longlong(v + FIX_MAGIC) - FIX_MAGIC_BITS;
    }
//! Scale applied to a value before it enters the fixed-point accumulator: the step's scale
//! fx[0], written by the tile scan.
template<typename R> __device__ __forceinline__ R acc_scale(const float* fx)
    {
    return R(fx[0]);
    }
//! LDS accumulation of a value already multiplied by acc_scale<R>().
__device__ __forceinline__ void lds_add_scaled(long long* p, double v)
    {
    __hip_atomic_fetch_add(p, round_to_ll(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
__device__ __forceinline__ void lds_add_scaled(int* p, float v)
    {
    __hip_atomic_fetch_add(p, round_to_int(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
__device__ __forceinline__ void glob_add(unsigned long long* p, long long v)
    {
    atomicAdd(p, static_cast<unsigned long long>(v));
    }
// ---------------------------------------------------------------------------
// Split-table lookup (EvaluatorPairPSWF.h:113-143, reduced s = r / rc).
// Record per segment: r_lo, dr_inv, S c0..c4, dS/ds c0..c4 (12 reals).
// ---------------------------------------------------------------------------
template<typename R>
__device__ __forceinline__ void split_lookup(const R* __restrict__ tab, unsigned int n_segs, R r, R rc, R inv_rc, R& L, R& negdLdr)
    {
    if (r < R(1.0e-12) || r > rc)
        {
        L = R(0);
        negdLdr = R(0);
        return;
        }
    // segment from the clamped reduced distance, offset t from the unclamped one
    const R su = r * inv_rc;
    R s = su < R(0) ? R(0) : (su > R(1) - R(1.0e-9) ? R(1) - R(1.0e-9) : su);
    unsigned int seg = static_cast<unsigned int>(s * R(n_segs));
    seg = seg >= n_segs ? n_segs - 1u : seg;
    const R* c = tab + size_t(seg) * 12u;
    const R t = (su - c[0]) * c[1];
    R S = c[6];
    S = c[5] + t * S;
    S = c[4] + t * S;
    S = c[3] + t * S;
    S = c[2] + t * S;
    R dS = c[11];
    dS = c[10] + t * dS;
    dS = c[9] + t * dS;
    dS = c[8] + t * dS;
    dS = c[7] + t * dS;
    dS *= inv_rc;
    const R inv_4pi_r = R(0.079577471545947673) / r;
    const R inv_4pi_r2 = inv_4pi_r / r;
    L = S * inv_4pi_r;
    negdLdr = S * inv_4pi_r2 - dS * inv_4pi_r;
    }

    } // namespace dev
    } // namespace esp
