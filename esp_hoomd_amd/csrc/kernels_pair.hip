// Short-range ESP correction over a HOOMD-layout full neighbour list, with the excluded-pair
// correction fused in.  Each particle is handled by a group of G lanes; the ten per-particle
// sums (force, energy, virial) are reduced with cross-lane shuffles inside the wave.
//
// Two evaluations of the split function S(r / r_c):
//  * compact (default): a 64-segment, degree-11 polynomial table of S held in LDS; dS/ds is
//    the analytic derivative of the same polynomial, so the pair force is exactly the
//    gradient of the pair energy.  The grid is persistent so each block loads the table once.
//  * reference format: the n_segs x 5 Horner tables of PSWF_Coeffs.h read from global memory
//    with the lookup of EvaluatorPairPSWF.h:113-143 (used when an external table is set).
//
// Reference counterparts: EvaluatorPairPSWF::evalForceAndEnergy (EvaluatorPairPSWF.h:274-295)
// called by HOOMD's PotentialPairGPU, and gpu_esp_fix_exclusions_kernel
// (ESPForceComputeGPU.cu:391-471).
#include "esp_device.h"
#include "esp_kernels.h"

#include <algorithm>
#include <cstdlib>

namespace esp
{
namespace dev
{

__device__ __forceinline__ double min_image(double d, double L, double invL)
    {
    return d - L * rint(d * invL);
    }

//! Minimum image of d = r_i - r_j in a tilted box, HOOMD's BoxDim::minImage order: lattice
//! vector a3 by rint(d_z / Lz), then a2 by rint(d_y / Ly), then a1 by rint(d_x / Lx).  For
//! |d| below half the smallest perpendicular width of the box this is the nearest image.
//! Explicit fused multiply-adds: single-system and batched kernels round alike.
__device__ __forceinline__ void min_image_tilted(double& dx, double& dy, double& dz, const kernel::PairArgs& a)
    {
    const double iz = rint(dz * (1.0 / a.Lz));
    dz = fma(-a.Lz, iz, dz);
    dy = fma(-(a.yz * a.Lz), iz, dy);
    dx = fma(-(a.xz * a.Lz), iz, dx);
    const double iy = rint(dy * (1.0 / a.Ly));
    dy = fma(-a.Ly, iy, dy);
    dx = fma(-(a.xy * a.Ly), iy, dx);
    dx = fma(-a.Lx, rint(dx * (1.0 / a.Lx)), dx);
    }

// Neighbour records of the pair pass, written by the pack kernel.
//  * FP64: (x, y, z, q) as doubles, 32 bytes; minimum image by rounding.
//  * FP32 (mixed): x, y, z as 32-bit fixed-point fractions of the box edge and q as a float,
//    16 bytes, so each neighbour costs one 16-byte load.  The difference of two fractions
//    taken modulo 2^32 and read as a signed integer is the minimum-image separation in units
//    of L / 2^32 (5e-8 at L = 215, far below FP32 resolution of the separation itself).
template<typename R> struct PairRec;
template<> struct PairRec<double>
    {
    using type = double4;
    };
template<> struct PairRec<float>
    {
    using type = uint4;
    };

//! Fixed-point fraction of the periodic box edge, x in box coordinates centred on the origin.
__device__ __forceinline__ unsigned int box_fraction(double x, double invL)
    {
    double u = x * invL + 0.5;
    u -= floor(u);
    return static_cast<unsigned int>(static_cast<unsigned long long>(u * 4294967296.0));
    }

template<typename R> __device__ __forceinline__ double rec_charge(const typename PairRec<R>::type& p)
    {
    if constexpr (sizeof(R) == 8)
        return p.w;
    else
        return double(__uint_as_float(p.w));
    }

//! Minimum-image separation r_i - r_j of two records and the charge product.  TILT selects the
//! tilted-box form at compile time, so the orthorhombic kernels carry none of its registers.
template<typename R, bool TILT>
__device__ __forceinline__ void rec_separation(const typename PairRec<R>::type& pi,
                                               const typename PairRec<R>::type& pj,
                                               const kernel::PairArgs& a,
                                               R& dx,
                                               R& dy,
                                               R& dz,
                                               R& qiqj)
    {
    if constexpr (sizeof(R) == 8)
        {
        if constexpr (TILT)
            {
            dx = pi.x - pj.x;
            dy = pi.y - pj.y;
            dz = pi.z - pj.z;
            min_image_tilted(dx, dy, dz, a);
            }
        else
            {
            dx = min_image(pi.x - pj.x, a.Lx, 1.0 / a.Lx);
            dy = min_image(pi.y - pj.y, a.Ly, 1.0 / a.Ly);
            dz = min_image(pi.z - pj.z, a.Lz, 1.0 / a.Lz);
            }
        qiqj = pi.w * pj.w;
        }
    else
        {
        constexpr double TWO_M32 = 2.3283064365386962890625e-10;  // 2^-32
        const float ux = float(int(pi.x - pj.x)), uy = float(int(pi.y - pj.y)), uz = float(int(pi.z - pj.z));
        dx = ux * float(a.Lx * TWO_M32);
        dy = uy * float(a.Ly * TWO_M32);
        dz = uz * float(a.Lz * TWO_M32);
        if constexpr (TILT)
            {
            // d = B (u_i - u_j): the wrapped fraction differences through the lattice vectors
            dx = __fmaf_rn(uz, float(a.xz * a.Lz * TWO_M32), __fmaf_rn(uy, float(a.xy * a.Ly * TWO_M32), dx));
            dy = __fmaf_rn(uz, float(a.yz * a.Lz * TWO_M32), dy);
            }
        qiqj = __uint_as_float(pi.w) * __uint_as_float(pj.w);
        }
    }

//! S and dS/ds from the compact table: segment = floor(s * NSEG), z = 2 t - 1.  The degree is
//! fixed at compile time so the row is fetched with 16-byte LDS loads and Horner is unrolled.
template<typename R>
__device__ __forceinline__ void compact_split(const R* __restrict__ tab, int nseg, R s, R& S, R& dSdz)
    {
    constexpr int NC = int(kernel::pair_compact_deg<R>()) + 1;
    constexpr int PER16 = 16 / int(sizeof(R));
    static_assert(NC % PER16 == 0, "compact rows must be whole 16-byte words");
    const R u = s * R(nseg);
    // s >= 0, so only the top end (s = 1, or s a rounding step above it) needs the clamp
    const int seg = min(int(u), nseg - 1);
    const R z = R(2) * (u - R(seg)) - R(1);
    R c[NC];
    const uint4* row = reinterpret_cast<const uint4*>(tab + seg * NC);
#pragma unroll
    for (int w = 0; w < NC / PER16; ++w)
        {
        const uint4 x = row[w];
        const unsigned int bits[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
        for (int k = 0; k < PER16; ++k)
            {
            if constexpr (sizeof(R) == 4)
                c[w * PER16 + k] = __uint_as_float(bits[k]);
            else
                c[w * PER16 + k] = __hiloint2double(int(bits[2 * k + 1]), int(bits[2 * k]));
            }
        }
    // value and derivative by one Horner pass (its first derivative step, 0 * z + c[NC-1], peeled)
    R d = c[NC - 1];
    R v = d * z + c[NC - 2];
#pragma unroll
    for (int j = NC - 3; j >= 0; --j)
        {
        d = d * z + v;
        v = v * z + c[j];
        }
    S = v;
    dSdz = d;  // dS/ds = dS/dz * 2 nseg
    }

//! r and 1/r from r^2: exact square root and division in FP64; the hardware reciprocal square
//! root (1 ulp) in FP32.
template<typename R>
__device__ __forceinline__ void r_and_rinv(R rsq, R& r, R& rinv)
    {
    if constexpr (sizeof(R) == 4)
        {
        // v_rsq_f32 without the denormal-input scaling (callers pass rsq >= 1e-24)
        rinv = __builtin_amdgcn_rsqf(rsq);
        r = rsq * rinv;
        }
    else
        {
        r = sqrt(rsq);
        rinv = R(1) / r;
        }
    }

//! L(r) and (-dL/dr) / r given rinv = 1 / r.  The compact form expects 1e-12 <= r and
//! r^2 < rc^2 (r may round to an ulp above rc: the last segment's polynomial is then taken a
//! rounding step past its end), which its callers test; the segment-table form returns zero
//! outside (1e-12, rc].
template<typename R, bool COMPACT>
__device__ __forceinline__ void pair_terms(const R* __restrict__ tab,
                                           const kernel::PairArgs& a,
                                           R r,
                                           R rinv,
                                           R rc,
                                           R inv_rc,
                                           R& L,
                                           R& ndr)
    {
    if constexpr (COMPACT)
        {
        R S, dSdz;
        compact_split<R>(tab, int(a.n_compact), r * inv_rc, S, dSdz);
        const R dz_dr = R(2 * int(a.n_compact)) * inv_rc;  // loop invariant
        const R inv4pir = R(0.079577471545947673) * rinv;
        L = S * inv4pir;
        ndr = (S * rinv - dSdz * dz_dr) * inv4pir * rinv;
        }
    else
        {
        R nd;
        split_lookup<R>(tab, a.n_segs, r, rc, inv_rc, L, nd);
        ndr = nd * rinv;
        }
    }

//! Pair correction of particle i by the G lanes of its group (`sub` = lane within the group);
//! every lane of the group must call it.  Writes force (fx, fy, fz, energy) and, when
//! requested, the six per-particle virial components.
//! Neighbour records in flight per lane: the pass is bound by the latency of these gathers.
//! FP32 (16-byte records, 61 VGPRs at two): eight, which took the pass from 0.48 to 0.40 ms
//! alone and the 1e6-particle mixed step from 0.965 to 0.92 ms (sixteen: 1.13 ms beside the
//! mesh chain).  FP64 (32-byte records, 124 VGPRs): two; four gained nothing.  The batched
//! kernel keeps two (its many small systems have short lists: eight cost 4 % on the 512-system
//! batch); the order of accumulation, and so the result, does not depend on it.
template<typename R> constexpr int pair_inflight()
    {
    return sizeof(R) == 4 ? 8 : 2;
    }

//! Records in flight per lane in the batched pair kernel: two.  Measured on MI355X (mixed
//! precision): 512 systems of 500-50,000 particles 7.94 ms per step at two, 8.11 at four, 8.61
//! at eight; the served stream's capacity 4.4e8 particle-steps/s at two, 3.9e8 at four, 4.2e8
//! at eight (its lists are short and its blocks share the GPU with the mesh stages).
template<typename R> constexpr int batch_pair_inflight()
    {
    return 2;
    }

template<typename R, int G, bool COMPACT, int NF = pair_inflight<R>(), bool TILT = false>
__device__ __forceinline__ void pair_particle(unsigned int i, int sub, const kernel::PairArgs& a, const R* __restrict__ tab)
    {
    const R rc = R(a.rc);
    const R inv_rc = R(1.0 / a.rc);
    const R rc2 = rc * rc;
    R f[3] = {R(0), R(0), R(0)};
    R e = R(0);
    R v[6] = {R(0), R(0), R(0), R(0), R(0), R(0)};
    using Rec = typename PairRec<R>::type;
    const Rec* __restrict__ pq = reinterpret_cast<const Rec*>(a.pq);
    const Rec pi = pq[i];
    const double qi = rec_charge<R>(pi);
    // force += fdivr * d; energy and virial are accumulated doubled and halved at the end
    auto accumulate = [&](R fdivr, R dE, R dx, R dy, R dz)
    {
        const R tx = fdivr * dx, ty = fdivr * dy, tz = fdivr * dz;
        f[0] += tx;
        f[1] += ty;
        f[2] += tz;
        e += dE;
        v[0] += tx * dx;
        v[1] += tx * dy;
        v[2] += tx * dz;
        v[3] += ty * dy;
        v[4] += ty * dz;
        v[5] += tz * dz;
    };
    if (a.nlist != nullptr && qi != 0.0 && a.rc >= 1.0e-12)
        {
        auto add = [&](const Rec& pj)
            {
            R dx, dy, dz, qiqj;
            rec_separation<R, TILT>(pi, pj, a, dx, dy, dz, qiqj);
            const R rsq = dx * dx + dy * dy + dz * dz;
            // outside the cutoff, uncharged, or coincident (r < 1e-12: no defined force)
            if (!(rsq < rc2 && rsq >= R(1.0e-24)) || qiqj == R(0))
                return;
            R r, rinv;
            r_and_rinv<R>(rsq, r, rinv);
            R Lr, ndr;
            pair_terms<R, COMPACT>(tab, a, r, rinv, rc, inv_rc, Lr, ndr);
            accumulate(qiqj * ndr, qiqj * Lr, dx, dy, dz);
            };
        const unsigned int nn = a.n_neigh[i];
        const uint64_t head = a.head_list[i];
        // NF neighbours in flight per lane, and the list entries of the next NF are fetched
        // before these are evaluated.  Entries past the end are read at the clamped index
        // nn - 1 (in bounds) and not evaluated.  Neighbours are accumulated in list order, so
        // the sums are the same as one at a time.
        if (unsigned(sub) < nn)
            {
            const unsigned int last = nn - 1u;
            const unsigned int* __restrict__ nl = a.nlist + head;
            unsigned int jn[NF];
#pragma unroll
            for (int f = 0; f < NF; ++f)
                jn[f] = nl[min(unsigned(sub) + unsigned(f * G), last)];
            for (unsigned int k = sub; k < nn; k += NF * G)
                {
                Rec pf[NF];
#pragma unroll
                for (int f = 0; f < NF; ++f)
                    pf[f] = pq[jn[f]];
                const unsigned int kn = k + NF * G;
#pragma unroll
                for (int f = 0; f < NF; ++f)
                    jn[f] = nl[min(kn + unsigned(f * G), last)];
                add(pf[0]);
#pragma unroll
                for (int f = 1; f < NF; ++f)
                    if (k + unsigned(f * G) < nn)
                        add(pf[f]);
                }
            }
        }
    if (a.n_ex != nullptr && qi != 0.0)
        {
        // remove the mesh part M(r) = 1/(4 pi r) - L(r) of each excluded pair at any distance
        const unsigned int ne = a.n_ex[i];
        for (unsigned int k = sub; k < ne; k += G)
            {
            const unsigned int j = a.ex_list[size_t(i) * a.ex_pitch + k];
            if (j >= a.N)
                continue;
            const Rec pj = pq[j];
            R dx, dy, dz, qiqj;
            rec_separation<R, TILT>(pi, pj, a, dx, dy, dz, qiqj);
            const R rsq = dx * dx + dy * dy + dz * dz;
            if (rsq < R(1.0e-20))
                continue;
            R r, rinv;
            r_and_rinv<R>(rsq, r, rinv);
            R Lr = R(0), ndr = R(0);
            if (rsq < rc2)
                pair_terms<R, COMPACT>(tab, a, r, rinv, rc, inv_rc, Lr, ndr);
            const R inv4pir = R(0.079577471545947673) * rinv;
            accumulate(qiqj * (ndr - inv4pir * rinv * rinv), qiqj * (Lr - inv4pir), dx, dy, dz);
            }
        }
    e *= R(0.5);
#pragma unroll
    for (int c = 0; c < 6; ++c)
        v[c] *= R(0.5);
    // reduce over the G lanes of this particle (every lane of the group takes part)
#pragma unroll
    for (int c = 0; c < 3; ++c)
        f[c] = group_sum<G>(f[c]);
    e = group_sum<G>(e);
#pragma unroll
    for (int c = 0; c < 6; ++c)
        v[c] = group_sum<G>(v[c]);
    if (sub == 0)
        {
        double4 o;
        o.x = double(f[0]);
        o.y = double(f[1]);
        o.z = double(f[2]);
        o.w = double(e);
        reinterpret_cast<double4*>(a.force)[i] = o;
        if (a.virial != nullptr)
            {
#pragma unroll
            for (int c = 0; c < 6; ++c)
                a.virial[size_t(c) * a.virial_pitch + i] = double(v[c]);
            }
        }
    }

//! Copy the compact split table of one system into LDS (callers synchronise around it).
template<typename R>
__device__ __forceinline__ const R* stage_compact_table(const kernel::PairArgs& a, unsigned char* smem)
    {
    R* t = reinterpret_cast<R*>(smem);
    const R* src = reinterpret_cast<const R*>(a.compact_tab);
    const int n = int(a.n_compact * (a.compact_deg + 1));
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        t[i] = src[i];
    return t;
    }

//! Single-system pair kernel, persistent grid.  Blocks are placed on the 8 XCDs round-robin
//! (block b on XCD (b + k) mod 8 for a launch-dependent k), so the blocks with one value of
//! b mod 8 share an XCD: they take one contiguous eighth of the particles and sweep it
//! together (interleaved), so that XCD's L2 holds the neighbour records of one moving window.
template<typename R, int G, bool COMPACT, int NF, bool TILT>
__global__ void __launch_bounds__(256) k_pair(kernel::PairArgs a)
    {
    extern __shared__ __align__(16) unsigned char smem[];
    const R* tab = reinterpret_cast<const R*>(a.split_tab);
    if constexpr (COMPACT)
        {
        tab = stage_compact_table<R>(a, smem);
        __syncthreads();
        }
    const int sub = int(threadIdx.x % G);
    const unsigned int per_pass = blockDim.x / G;
    const unsigned int nb = gridDim.x;
    const unsigned int nx = (nb % 8u == 0u) ? 8u : 1u;
    const unsigned int x = blockIdx.x % nx, j = blockIdx.x / nx, per_x = nb / nx;
    const unsigned int span = (a.N + nx - 1u) / nx;
    const unsigned int p0 = x * span;
    const unsigned int p1 = min(a.N, p0 + span);
    for (unsigned int i = p0 + j * per_pass + threadIdx.x / G; i < p1; i += per_x * per_pass)
        pair_particle<R, G, COMPACT, NF, TILT>(i, sub, a, tab);
    }

//! Batched pair correction: items are (system, chunk of 256 / G particles).  The compact
//! table is restaged whenever the block moves to another system.  Systems whose flags
//! exclude the pair part get zero force and virial (the outputs are always overwritten).
template<typename R, int G, bool TILT>
__global__ void __launch_bounds__(256) kb_pair(const kernel::BatchSys* __restrict__ ds, kernel::BatchList l)
    {
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned int staged = 0xffffffffu;
    const int sub = int(threadIdx.x % G);
    for (unsigned int it = blockIdx.x; it < l.n; it += gridDim.x)
        {
        const kernel::BatchItem w = l.items[it];
        const kernel::PairArgs& a = ds[w.sys].pair;
        const bool compact = a.compact_tab != nullptr;
        if (compact && staged != w.sys)
            {
            __syncthreads();
            stage_compact_table<R>(a, smem);
            __syncthreads();
            staged = w.sys;
            }
        const unsigned int i = w.local * (blockDim.x / G) + threadIdx.x / G;
        if (i >= a.N)
            continue;
        if ((ds[w.sys].flags & 1u) == 0u)
            {
            if (sub == 0)
                {
                reinterpret_cast<double4*>(a.force)[i] = make_double4(0.0, 0.0, 0.0, 0.0);
                if (a.virial != nullptr)
                    for (int c = 0; c < 6; ++c)
                        a.virial[size_t(c) * a.virial_pitch + i] = 0.0;
                }
            continue;
            }
        // the tilted form only for tilted systems: an orthorhombic one keeps the bits it has alone
        if (TILT && a.tilted)
            {
            if (compact)
                pair_particle<R, G, true, batch_pair_inflight<R>(), true>(i, sub, a, reinterpret_cast<const R*>(smem));
            else
                pair_particle<R, G, false, batch_pair_inflight<R>(), true>(i, sub, a, reinterpret_cast<const R*>(a.split_tab));
            }
        else if (compact)
            pair_particle<R, G, true, batch_pair_inflight<R>(), false>(i, sub, a, reinterpret_cast<const R*>(smem));
        else
            pair_particle<R, G, false, batch_pair_inflight<R>(), false>(i, sub, a, reinterpret_cast<const R*>(a.split_tab));
        }
    }

template<typename R> __device__ __forceinline__ void pack_one(unsigned int i, const kernel::PairArgs& a)
    {
    const double4 p = reinterpret_cast<const double4*>(a.postype)[i];
    if constexpr (sizeof(R) == 8)
        reinterpret_cast<double4*>(a.pq)[i] = make_double4(p.x, p.y, p.z, a.charge[i]);
    else if (!a.tilted)
        reinterpret_cast<uint4*>(a.pq)[i] = make_uint4(box_fraction(p.x, 1.0 / a.Lx), box_fraction(p.y, 1.0 / a.Ly),
                                                      box_fraction(p.z, 1.0 / a.Lz), __float_as_uint(float(a.charge[i])));
    else
        {
        // fractions along the lattice vectors (HOOMD's makeFraction), B^-1 r
        const double fx = fma(-fma(-a.xy, a.yz, a.xz), p.z, fma(-a.xy, p.y, p.x)) * (1.0 / a.Lx);
        const double fy = fma(-a.yz, p.z, p.y) * (1.0 / a.Ly);
        const double fz = p.z * (1.0 / a.Lz);
        reinterpret_cast<uint4*>(a.pq)[i] = make_uint4(box_fraction(fx, 1.0), box_fraction(fy, 1.0), box_fraction(fz, 1.0),
                                                      __float_as_uint(float(a.charge[i])));
        }
    }

template<typename R> __global__ void __launch_bounds__(256) k_pack_pq(kernel::PairArgs a)
    {
    const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < a.N)
        pack_one<R>(i, a);
    }

//! Batched packing: items are (system, chunk of 256 particles).
template<typename R> __global__ void __launch_bounds__(256) kb_pack_pq(const kernel::BatchSys* __restrict__ ds, kernel::BatchList l)
    {
    for (unsigned int it = blockIdx.x; it < l.n; it += gridDim.x)
        {
        const kernel::BatchItem w = l.items[it];
        const kernel::PairArgs& a = ds[w.sys].pair;
        const unsigned int i = w.local * 256u + threadIdx.x;
        if (i < a.N)
            pack_one<R>(i, a);
        }
    }

    } // namespace dev

namespace kernel
{
template<typename R, int G> static void launch_pair(const PairArgs& a, unsigned int nb, hipStream_t s)
    {
    constexpr int NF = dev::pair_inflight<R>();
    const size_t lds = size_t(a.n_compact) * (a.compact_deg + 1) * sizeof(R);
    if (a.compact_tab != nullptr)
        {
        if (a.tilted)
            hipLaunchKernelGGL((dev::k_pair<R, G, true, NF, true>), dim3(nb), dim3(256), lds, s, a);
        else
            hipLaunchKernelGGL((dev::k_pair<R, G, true, NF, false>), dim3(nb), dim3(256), lds, s, a);
        }
    else if (a.tilted)
        hipLaunchKernelGGL((dev::k_pair<R, G, false, NF, true>), dim3(nb), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((dev::k_pair<R, G, false, NF, false>), dim3(nb), dim3(256), 0, s, a);
    }

template<typename R> hipError_t pack_positions_charges(const PairArgs& a, hipStream_t s)
    {
    if (a.N == 0)
        return hipSuccess;
    hipLaunchKernelGGL(dev::k_pack_pq<R>, dim3((a.N + 255u) / 256u), dim3(256), 0, s, a);
    return hipGetLastError();
    }
template hipError_t pack_positions_charges<double>(const PairArgs&, hipStream_t);
template hipError_t pack_positions_charges<float>(const PairArgs&, hipStream_t);

template<typename R> hipError_t batch_pack(const BatchSys* d, BatchList l, hipStream_t s)
    {
    if (l.n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(dev::kb_pack_pq<R>, dim3(l.grid), dim3(256), 0, s, d, l);
    return hipGetLastError();
    }
template hipError_t batch_pack<double>(const BatchSys*, BatchList, hipStream_t);
template hipError_t batch_pack<float>(const BatchSys*, BatchList, hipStream_t);

template<typename R> hipError_t pair_forces(const PairArgs& a, int lanes_per_particle, unsigned int max_blocks, hipStream_t s)
    {
    if (a.N == 0)
        return hipSuccess;
    const size_t threads = size_t(a.N) * lanes_per_particle;
    // persistent grid: by default at most 8 blocks per CU of a 256-CU device; the caller caps
    // it lower when the pass shares the GPU with the mesh chain (environment
    // ESP_PAIR_MAX_BLOCKS overrides both)
    static int env_blocks = -2;
    if (env_blocks == -2)
        {
        const char* e = getenv("ESP_PAIR_MAX_BLOCKS");
        env_blocks = e ? std::max(8, atoi(e)) : -1;
        }
    const size_t cap = env_blocks > 0 ? size_t(env_blocks) : (max_blocks ? size_t(max_blocks) : size_t(2048));
    unsigned int nb = unsigned(std::min<size_t>((threads + 255) / 256, cap));
    if (nb >= 8)
        nb &= ~7u;  // a multiple of the 8 XCDs for the XCD-aware mapping
    switch (lanes_per_particle)
        {
        case 1: launch_pair<R, 1>(a, nb, s); break;
        case 2: launch_pair<R, 2>(a, nb, s); break;
        case 4: launch_pair<R, 4>(a, nb, s); break;
        case 8: launch_pair<R, 8>(a, nb, s); break;
        case 16: launch_pair<R, 16>(a, nb, s); break;
        case 32: launch_pair<R, 32>(a, nb, s); break;
        default: return hipErrorInvalidValue;
        }
    return hipGetLastError();
    }
template hipError_t pair_forces<double>(const PairArgs&, int, unsigned int, hipStream_t);
template hipError_t pair_forces<float>(const PairArgs&, int, unsigned int, hipStream_t);

template<typename R> hipError_t batch_pair(const BatchSys* d, BatchList l, size_t lds, bool any_tilted, hipStream_t s)
    {
    if (l.n == 0)
        return hipSuccess;
    // the tilted-box instantiation only when some system of the batch is tilted (it handles
    // orthorhombic systems too, as the special case of zero tilt)
    if (any_tilted)
        hipLaunchKernelGGL((dev::kb_pair<R, BATCH_PAIR_LANES, true>), dim3(l.grid), dim3(256), lds, s, d, l);
    else
        hipLaunchKernelGGL((dev::kb_pair<R, BATCH_PAIR_LANES, false>), dim3(l.grid), dim3(256), lds, s, d, l);
    return hipGetLastError();
    }
template hipError_t batch_pair<double>(const BatchSys*, BatchList, size_t, bool, hipStream_t);
template hipError_t batch_pair<float>(const BatchSys*, BatchList, size_t, bool, hipStream_t);
    } // namespace kernel
    } // namespace esp
