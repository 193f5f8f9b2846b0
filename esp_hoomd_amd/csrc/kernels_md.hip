// Molecular-dynamics kernels around the ESP hot path (the roles HOOMD plays for the reference
// plugin, ESPForceCompute.h:39,177): cell-list neighbour lists in the HOOMD full-list layout,
// the WCA repulsion, velocity-Verlet NVE integration, a Langevin thermostat and the
// displacement check that triggers neighbour-list rebuilds.
#include "esp_device.h"
#include "esp_md.h"

#include <algorithm>

namespace esp
{
namespace dev
{

__device__ __forceinline__ double wrap_min_image(double d, double L, double invL)
    {
    return d - L * rint(d * invL);
    }

__device__ __forceinline__ int cell_coord(double x, double L, int n)
    {
    double f = x / L + 0.5;
    f -= floor(f);
    int c = int(f * n);
    return c >= n ? n - 1 : (c < 0 ? 0 : c);
    }

__global__ void __launch_bounds__(256) k_cell_count(unsigned int N, const double4* __restrict__ pos, md::CellGeom cg,
                                                    unsigned int* __restrict__ count, unsigned int* __restrict__ cell,
                                                    unsigned int* __restrict__ slot)
    {
    const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    const double4 r = pos[i];
    const int cx = cell_coord(r.x, cg.Lx, cg.nx), cy = cell_coord(r.y, cg.Ly, cg.ny), cz = cell_coord(r.z, cg.Lz, cg.nz);
    const unsigned int c = unsigned((cz * cg.ny + cy) * cg.nx + cx);
    cell[i] = c;
    slot[i] = atomicAdd(&count[c], 1u);
    }

//! Cell members in arrival order (the slots come from atomics, so the order varies by run).
__global__ void __launch_bounds__(256) k_cell_fill(unsigned int N, const unsigned int* __restrict__ cell, const unsigned int* __restrict__ slot,
                                                   const unsigned int* __restrict__ start, unsigned int* __restrict__ members)
    {
    const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    members[start[cell[i]] + slot[i]] = i;
    }

//! Wrap a coordinate into [-L/2, L/2) the way cell_coord does.
__device__ __forceinline__ double wrap_into_box(double x, double L)
    {
    double f = x / L + 0.5;
    f -= floor(f);
    return (f - 0.5) * L;
    }

//! Cell members in ascending particle index: particle i's rank in its cell is the number of
//! members below i.  The neighbour rows then list candidates in an order fixed by the
//! positions alone, so pair sums over them are reproducible bit for bit.  Each member's
//! position, wrapped into the box, is stored in the same slot (`spos`), so the list pass
//! reads every cell's candidates as one contiguous run.
__global__ void __launch_bounds__(256) k_cell_order(unsigned int N, const double4* __restrict__ pos, md::CellGeom cg,
                                                    const unsigned int* __restrict__ cell, const unsigned int* __restrict__ start,
                                                    const unsigned int* __restrict__ unsorted, unsigned int* __restrict__ members,
                                                    double4* __restrict__ spos)
    {
    const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    const unsigned int c = cell[i];
    const unsigned int m0 = start[c], m1 = start[c + 1];
    unsigned int rank = 0;
    for (unsigned int m = m0; m < m1; ++m)
        rank += unsorted[m] < i ? 1u : 0u;
    members[m0 + rank] = i;
    const double4 r = pos[i];
    spos[m0 + rank] = make_double4(wrap_into_box(r.x, cg.Lx), wrap_into_box(r.y, cg.Ly), wrap_into_box(r.z, cg.Lz), r.w);
    }

//! Full neighbour list of one particle (r < r_list, exclusions left out) by the G lanes of its
//! group.  Groups take particles in cell-sorted order (slot m of the member list), so the
//! groups of a wave share their 27 neighbour cells and the candidates' positions -- read from
//! the cell-sorted copy `spos`, one contiguous run per cell -- stay in cache.  Each lane tests
//! every G-th member of a cell, and the accepted candidates are compacted with a ballot into
//! consecutive slots of the particle's row (row i, HOOMD layout), in cell-walk order and, within
//! a cell, in ascending index.  Along an axis with at least 3 cells the periodic image of a
//! neighbour cell is fixed by the walk (a shift of -L or +L where the cell index wraps), so no
//! minimum-image rounding is needed per candidate; axes with fewer cells use rint.  With
//! r_inner2 > 0 the same pass also writes the inner list (the entries with r^2 < r_inner2) to
//! rows of width nmax_in; overflow[0] / overflow[1] collect the largest full / inner counts
//! that did not fit.  Every lane of the group must call it; `active` (m < N) is the same for
//! the whole group.
constexpr int NLIST_LANES = 8;
template<int G>
__device__ __forceinline__ void nlist_group(unsigned int m, int sub, bool active, const double4* __restrict__ spos,
                                            const md::CellGeom& cg, const unsigned int* __restrict__ cell,
                                            const unsigned int* __restrict__ start, const unsigned int* __restrict__ members,
                                            double r_list2, const unsigned int* __restrict__ n_ex,
                                            const unsigned int* __restrict__ ex_list, unsigned int ex_pitch, unsigned int nmax,
                                            unsigned int* __restrict__ n_neigh, unsigned int* __restrict__ nlist,
                                            uint64_t* __restrict__ head, unsigned int* __restrict__ overflow, const md::InnerList& in)
    {
    if (!active)
        return;
    const unsigned int i = members[m];
    const double4 ri = spos[m];
    const unsigned int ci = cell[i];
    const int cx = int(ci % unsigned(cg.nx)), cy = int((ci / unsigned(cg.nx)) % unsigned(cg.ny)), cz = int(ci / unsigned(cg.nx * cg.ny));
    const double invLx = 1.0 / cg.Lx, invLy = 1.0 / cg.Ly, invLz = 1.0 / cg.Lz;
    const bool shx = cg.nx >= 3, shy = cg.ny >= 3, shz = cg.nz >= 3;
    const unsigned int ne = n_ex ? n_ex[i] : 0u;
    const unsigned int lane = __lane_id();
    const unsigned int gshift = lane & ~unsigned(G - 1);         // first lane of the group
    const unsigned long long gmask = (G == 64) ? ~0ull : ((1ull << G) - 1ull);
    const unsigned long long below = (1ull << unsigned(sub)) - 1ull;
    unsigned int* row = nlist + size_t(i) * nmax;
    unsigned int* row_in = in.r2 > 0.0 ? in.list + size_t(i) * in.nmax : nullptr;
    unsigned int k = 0, k_in = 0;
    // candidates mm in [m0, m1) (cell-sorted, one image shift (bx, by, bz) = r_i - s for all of
    // them), compacted into the rows with one ballot per G candidates
    auto scan_run = [&](unsigned int m0, unsigned int m1, double bx, double by, double bz)
    {
        for (unsigned int mm0 = m0; mm0 < m1; mm0 += unsigned(G))
            {
            const unsigned int mm = mm0 + unsigned(sub);
            unsigned int j = 0;
            bool take = false;
            double r2 = 0.0;
            if (mm < m1 && mm != m)
                {
                const double4 rj = spos[mm];
                double ddx = bx - rj.x, ddy = by - rj.y, ddz = bz - rj.z;
                if (!shx)
                    ddx -= cg.Lx * rint(ddx * invLx);
                if (!shy)
                    ddy -= cg.Ly * rint(ddy * invLy);
                if (!shz)
                    ddz -= cg.Lz * rint(ddz * invLz);
                r2 = ddx * ddx + ddy * ddy + ddz * ddz;
                take = r2 < r_list2;
                if (take)
                    {
                    j = members[mm];
                    for (unsigned int e = 0; take && e < ne; ++e)
                        take = ex_list[size_t(i) * ex_pitch + e] != j;
                    }
                }
            const unsigned long long accepted = (__ballot(take) >> gshift) & gmask;
            const unsigned int slot = k + unsigned(__popcll(accepted & below));
            if (take && slot < nmax)
                row[slot] = j;
            k += unsigned(__popcll(accepted));
            if (in.r2 > 0.0)
                {
                const bool inner = take && r2 < in.r2;
                const unsigned long long acc_in = (__ballot(inner) >> gshift) & gmask;
                const unsigned int slot_in = k_in + unsigned(__popcll(acc_in & below));
                if (inner && slot_in < in.nmax)
                    row_in[slot_in] = j;
                k_in += unsigned(__popcll(acc_in));
                }
            }
    };
    if (cg.reach == 2)
        {
        // half-size cells, every axis with at least five: the 5 x 5 rows of five x-cells around
        // the particle; the x cells of a row are consecutive in cell order, so a row is one run
        // of candidates (two or three where it wraps, each with its own image shift)
        for (int dz = -2; dz <= 2; ++dz)
            for (int dy = -2; dy <= 2; ++dy)
                {
                const int uy = cy + dy, uz = cz + dz;
                const int y = (uy + cg.ny) % cg.ny, z = (uz + cg.nz) % cg.nz;
                const double sy = uy < 0 ? -cg.Ly : (uy >= cg.ny ? cg.Ly : 0.0);
                const double sz = uz < 0 ? -cg.Lz : (uz >= cg.nz ? cg.Lz : 0.0);
                const double by = ri.y - sy, bz = ri.z - sz;
                const unsigned int rowc = unsigned(z * cg.ny + y) * unsigned(cg.nx);
                const int lo = cx - 2, hi = cx + 2;
                scan_run(start[rowc + unsigned(max(lo, 0))], start[rowc + unsigned(min(hi, cg.nx - 1)) + 1], ri.x, by, bz);
                if (lo < 0)  // cells nx + lo .. nx - 1, seen one box length below
                    scan_run(start[rowc + unsigned(cg.nx + lo)], start[rowc + unsigned(cg.nx)], ri.x + cg.Lx, by, bz);
                if (hi >= cg.nx)  // cells 0 .. hi - nx, one box length above
                    scan_run(start[rowc], start[rowc + unsigned(hi - cg.nx) + 1], ri.x - cg.Lx, by, bz);
                }
        }
    else
        {
        // neighbour cells, each visited once even when an axis has fewer than 3 cells
        const int rx = cg.nx >= 2 ? 1 : 0, ry = cg.ny >= 2 ? 1 : 0, rz = cg.nz >= 2 ? 1 : 0;
        for (int dz = -rz; dz <= (cg.nz == 2 ? 0 : rz); ++dz)
            for (int dy = -ry; dy <= (cg.ny == 2 ? 0 : ry); ++dy)
                for (int dx = -rx; dx <= (cg.nx == 2 ? 0 : rx); ++dx)
                    {
                    const int ux = cx + dx, uy = cy + dy, uz = cz + dz;
                    const int x = (ux + cg.nx) % cg.nx, y = (uy + cg.ny) % cg.ny, z = (uz + cg.nz) % cg.nz;
                    // image of the neighbour cell next to particle i: r_j + s
                    const double sx = shx ? (ux < 0 ? -cg.Lx : (ux >= cg.nx ? cg.Lx : 0.0)) : 0.0;
                    const double sy = shy ? (uy < 0 ? -cg.Ly : (uy >= cg.ny ? cg.Ly : 0.0)) : 0.0;
                    const double sz = shz ? (uz < 0 ? -cg.Lz : (uz >= cg.nz ? cg.Lz : 0.0)) : 0.0;
                    const unsigned int c = unsigned((z * cg.ny + y) * cg.nx + x);
                    scan_run(start[c], start[c + 1], ri.x - sx, ri.y - sy, ri.z - sz);
                    }
        }
    if (sub == 0)
        {
        n_neigh[i] = k < nmax ? k : nmax;
        head[i] = uint64_t(i) * nmax;
        if (k > nmax)
            atomicMax(overflow, k);
        if (in.r2 > 0.0)
            {
            in.n[i] = k_in < in.nmax ? k_in : in.nmax;
            in.head[i] = uint64_t(i) * in.nmax;
            if (k_in > in.nmax)
                atomicMax(overflow + 1, k_in);
            }
        }
    }

__global__ void __launch_bounds__(256) k_nlist(unsigned int N, const double4* __restrict__ spos, md::CellGeom cg,
                                               const unsigned int* __restrict__ cell, const unsigned int* __restrict__ start,
                                               const unsigned int* __restrict__ members, double r_list2, const unsigned int* __restrict__ n_ex,
                                               const unsigned int* __restrict__ ex_list, unsigned int ex_pitch, unsigned int nmax,
                                               unsigned int* __restrict__ n_neigh, unsigned int* __restrict__ nlist,
                                               uint64_t* __restrict__ head, unsigned int* __restrict__ overflow, md::InnerList in)
    {
    const unsigned int t = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned int m = t / NLIST_LANES;
    nlist_group<NLIST_LANES>(m, int(t % NLIST_LANES), m < N, spos, cg, cell, start, members, r_list2, n_ex, ex_list, ex_pitch, nmax,
                             n_neigh, nlist, head, overflow, in);
    }

//! WCA force, energy and virial of particle i added to its outputs, by the G lanes of its group
//! (`sub` = lane within the group) over the (longer) ESP neighbour list: each lane takes every
//! G-th entry, so the group has G neighbour loads in flight (the list is mostly beyond the
//! WCA cutoff, so the pass is bound by those loads).  Every lane of the group must call it;
//! `active` (i < N) is the same for the whole group.
constexpr int WCA_LANES = 8;
template<int G>
__device__ __forceinline__ void wca_group(unsigned int i, int sub, bool active, const double4* __restrict__ pos, double Lx, double Ly,
                                          double Lz, const unsigned int* __restrict__ n_neigh, const unsigned int* __restrict__ nlist,
                                          const uint64_t* __restrict__ head, double sigma, double epsilon, double4* __restrict__ force,
                                          double* __restrict__ virial, size_t pitch)
    {
    double fx = 0, fy = 0, fz = 0, e = 0, v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
    if (active)
        {
        const double4 ri = pos[i];
        const double invLx = 1.0 / Lx, invLy = 1.0 / Ly, invLz = 1.0 / Lz;
        const double rc2 = 1.2599210498948732 * sigma * sigma;  // (2^(1/6) sigma)^2
        const double s2 = sigma * sigma;
        const unsigned int nn = n_neigh[i];
        const uint64_t h = head[i];
        for (unsigned int k = unsigned(sub); k < nn; k += unsigned(G))
            {
            const unsigned int j = nlist[h + k];
            const double4 rj = pos[j];
            const double dx = wrap_min_image(ri.x - rj.x, Lx, invLx);
            const double dy = wrap_min_image(ri.y - rj.y, Ly, invLy);
            const double dz = wrap_min_image(ri.z - rj.z, Lz, invLz);
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 >= rc2)
                continue;
            const double sr2 = s2 / r2;
            const double sr6 = sr2 * sr2 * sr2;
            const double fdivr = 24.0 * epsilon * sr6 * (2.0 * sr6 - 1.0) / r2;
            fx += fdivr * dx;
            fy += fdivr * dy;
            fz += fdivr * dz;
            e += 0.5 * (4.0 * epsilon * sr6 * (sr6 - 1.0) + epsilon);
            const double hf = 0.5 * fdivr;
            v0 += hf * dx * dx;
            v1 += hf * dx * dy;
            v2 += hf * dx * dz;
            v3 += hf * dy * dy;
            v4 += hf * dy * dz;
            v5 += hf * dz * dz;
            }
        }
    fx = group_sum<G>(fx);
    fy = group_sum<G>(fy);
    fz = group_sum<G>(fz);
    e = group_sum<G>(e);
    v0 = group_sum<G>(v0);
    v1 = group_sum<G>(v1);
    v2 = group_sum<G>(v2);
    v3 = group_sum<G>(v3);
    v4 = group_sum<G>(v4);
    v5 = group_sum<G>(v5);
    if (!active || sub != 0)
        return;
    double4 f = force[i];
    f.x += fx;
    f.y += fy;
    f.z += fz;
    f.w += e;
    force[i] = f;
    if (virial)
        {
        virial[0 * pitch + i] += v0;
        virial[1 * pitch + i] += v1;
        virial[2 * pitch + i] += v2;
        virial[3 * pitch + i] += v3;
        virial[4 * pitch + i] += v4;
        virial[5 * pitch + i] += v5;
        }
    }

//! WCA over a neighbour list of any length, one wave per 64 particles: when the wave's rows are
//! short (at most 16 entries, as in an inner list) each lane takes one particle in one pass;
//! otherwise the wave takes its 64 particles in WCA_LANES passes of groups of WCA_LANES lanes.
__global__ void __launch_bounds__(256) k_wca(unsigned int N, const double4* __restrict__ pos, double Lx, double Ly, double Lz,
                                             const unsigned int* __restrict__ n_neigh, const unsigned int* __restrict__ nlist,
                                             const uint64_t* __restrict__ head, double sigma, double epsilon, double4* __restrict__ force,
                                             double* __restrict__ virial, size_t pitch)
    {
    const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned int lane = threadIdx.x & 63u;
    const unsigned int longest = wave_max(i < N ? n_neigh[i] : 0u);
    if (longest <= 16u)
        {
        wca_group<1>(i, 0, i < N, pos, Lx, Ly, Lz, n_neigh, nlist, head, sigma, epsilon, force, virial, pitch);
        return;
        }
    const unsigned int first = i - lane;
    constexpr unsigned int PER_PASS = 64u / unsigned(WCA_LANES);
    for (unsigned int pass = 0; pass < unsigned(WCA_LANES); ++pass)
        {
        const unsigned int ip = first + pass * PER_PASS + lane / unsigned(WCA_LANES);
        wca_group<WCA_LANES>(ip, int(lane % unsigned(WCA_LANES)), ip < N, pos, Lx, Ly, Lz, n_neigh, nlist, head, sigma, epsilon, force,
                             virial, pitch);
        }
    }
//! Counter-based uniform random number in (-1, 1) from (seed, step, particle, component).
__device__ __forceinline__ double uniform_pm1(uint64_t seed, uint64_t step, uint64_t i, uint32_t c)
    {
    uint64_t x = seed * 0x9E3779B97F4A7C15ull ^ (step + 0x632BE59BD9B4E019ull) * 0xBF58476D1CE4E5B9ull;
    x ^= (i * 4 + c + 1) * 0x94D049BB133111EBull;
    // splitmix64 finaliser, twice
    for (int r = 0; r < 2; ++r)
        {
        x ^= x >> 30;
        x *= 0xBF58476D1CE4E5B9ull;
        x ^= x >> 27;
        x *= 0x94D049BB133111EBull;
        x ^= x >> 31;
        }
    return (double(x >> 11) + 0.5) * (2.0 / 9007199254740992.0) - 1.0;
    }

//! Velocity-Verlet first half for particle i: v += f/(2m) dt; x += v dt (wrapped into the box).
//! vel = (vx, vy, vz, mass).
//! First velocity-Verlet half: half kick v += (dt / 2m) f, drift r += dt v, wrap into the box.
//! Written with explicit fused multiply-adds, so the drift kernel and the predicted-displacement
//! check (kb_predict_disp) compute the same new position bit for bit.
__device__ __forceinline__ void vv_first_values(double4& r, double4& v, const double4& f, double dt, double Lx, double Ly, double Lz)
    {
    const double h = 0.5 * dt / v.w;
    v.x = fma(h, f.x, v.x);
    v.y = fma(h, f.y, v.y);
    v.z = fma(h, f.z, v.z);
    r.x = fma(v.x, dt, r.x);
    r.y = fma(v.y, dt, r.y);
    r.z = fma(v.z, dt, r.z);
    r.x = fma(-Lx, rint(r.x / Lx), r.x);
    r.y = fma(-Ly, rint(r.y / Ly), r.y);
    r.z = fma(-Lz, rint(r.z / Lz), r.z);
    }

__device__ __forceinline__ void vv_first_one(unsigned int i, double4* __restrict__ pos, double4* __restrict__ vel,
                                             const double4* __restrict__ force, double dt, double Lx, double Ly, double Lz)
    {
    double4 v = vel[i];
    double4 r = pos[i];
    vv_first_values(r, v, force[i], dt, Lx, Ly, Lz);
    pos[i] = r;
    vel[i] = v;
    }

//! Squared minimum-image distance between a position and its position at the last list build
//! (explicit fused multiply-adds: the same bits in every kernel that uses it).
__device__ __forceinline__ double displacement2(const double4& a, const double4& b, const double* L)
    {
    const double dx = fma(-L[0], rint((a.x - b.x) * (1.0 / L[0])), a.x - b.x);
    const double dy = fma(-L[1], rint((a.y - b.y) * (1.0 / L[1])), a.y - b.y);
    const double dz = fma(-L[2], rint((a.z - b.z) * (1.0 / L[2])), a.z - b.z);
    return fma(dx, dx, fma(dy, dy, dz * dz));
    }

__global__ void __launch_bounds__(256) k_vv_first(unsigned int N, double4* __restrict__ pos, double4* __restrict__ vel,
                                                  const double4* __restrict__ force, double dt, double Lx, double Ly, double Lz)
    {
    const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N)
        vv_first_one(i, pos, vel, force, dt, Lx, Ly, Lz);
    }

//! Velocity-Verlet second half for particle i with an optional Langevin bath (gamma > 0): the
//! drag -gamma v and a uniform random force of variance 2 gamma kT / dt per component are
//! added to f.
__device__ __forceinline__ void vv_second_one(unsigned int i, double4* __restrict__ vel, const double4* __restrict__ force, double dt,
                                              double gamma, double kT, uint64_t seed, uint64_t step)
    {
    double4 v = vel[i];
    const double4 f = force[i];
    const double h = 0.5 * dt / v.w;
    double bx = 0.0, by = 0.0, bz = 0.0;
    if (gamma > 0.0)
        {
        // drag and noise act over the whole step (one Euler-Maruyama kick per step), so that
        // the bath temperature is kT: per-step noise variance 2 gamma kT dt / m^2
        const double amp = sqrt(6.0 * gamma * kT / dt);  // uniform(-1,1) has variance 1/3
        bx = -gamma * v.x + amp * uniform_pm1(seed, step, i, 0);
        by = -gamma * v.y + amp * uniform_pm1(seed, step, i, 1);
        bz = -gamma * v.z + amp * uniform_pm1(seed, step, i, 2);
        }
    const double k = dt / v.w;
    v.x += h * f.x + k * bx;
    v.y += h * f.y + k * by;
    v.z += h * f.z + k * bz;
    vel[i] = v;
    }

__global__ void __launch_bounds__(256) k_vv_second(unsigned int N, double4* __restrict__ vel, const double4* __restrict__ force, double dt,
                                                   double gamma, double kT, uint64_t seed, uint64_t step)
    {
    const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N)
        vv_second_one(i, vel, force, dt, gamma, kT, seed, step);
    }

//! Raises *out (the bits of a non-negative double, monotone in unsigned comparison) to the
//! workgroup's largest d2: wave maxima through LDS, then one atomic per workgroup, skipped when
//! the value already there is at least as large (values only grow, so a stale read only costs
//! an unneeded atomic).  Every thread of the workgroup must call it.
__device__ __forceinline__ void block_max_to(double d2, unsigned long long* out)
    {
    __shared__ double wave_max[4];
    for (int off = 32; off > 0; off >>= 1)
        d2 = fmax(d2, __shfl_xor(d2, off, 64));
    if ((threadIdx.x & 63) == 0)
        wave_max[threadIdx.x >> 6] = d2;
    __syncthreads();
    if (threadIdx.x == 0)
        {
        for (unsigned int w = 1; w < (blockDim.x >> 6); ++w)
            d2 = fmax(d2, wave_max[w]);
        const unsigned long long bits = static_cast<unsigned long long>(__double_as_longlong(d2));
        if (bits > *static_cast<volatile unsigned long long*>(out))
            atomicMax(out, bits);
        }
    }

//! Largest squared minimum-image displacement since the last neighbour-list build, as the bits
//! of a non-negative double (monotone in unsigned comparison).
//! A grid of at most MAX_DISP_BLOCKS workgroups strides over the particles, so at most that many
//! atomics meet on the one result word (one per 256 particles serialised there for ~40 us at
//! 1e6 particles).
constexpr unsigned int MAX_DISP_BLOCKS = 512;
__global__ void __launch_bounds__(256) k_max_disp(unsigned int N, const double4* __restrict__ pos, const double4* __restrict__ last,
                                                  double Lx, double Ly, double Lz, unsigned long long* __restrict__ out)
    {
    const double invLx = 1.0 / Lx, invLy = 1.0 / Ly, invLz = 1.0 / Lz;
    double d2 = 0.0;
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        {
        const double4 a = pos[i], b = last[i];
        const double dx = wrap_min_image(a.x - b.x, Lx, invLx);
        const double dy = wrap_min_image(a.y - b.y, Ly, invLy);
        const double dz = wrap_min_image(a.z - b.z, Lz, invLz);
        d2 = fmax(d2, dx * dx + dy * dy + dz * dz);
        }
    block_max_to(d2, out);
    }

__global__ void __launch_bounds__(256) k_copy4(unsigned int N, const double4* __restrict__ src, double4* __restrict__ dst)
    {
    const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N)
        dst[i] = src[i];
    }

//! Kinetic energy and potential energy (sum of force.w) partial sums per block.
__global__ void __launch_bounds__(256) k_energies(unsigned int N, const double4* __restrict__ vel, const double4* __restrict__ force,
                                                  double* __restrict__ part)
    {
    const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    double ke = 0.0, pe = 0.0;
    if (i < N)
        {
        const double4 v = vel[i];
        ke = 0.5 * v.w * (v.x * v.x + v.y * v.y + v.z * v.z);
        pe = force[i].w;
        }
    ke = wave_sum(ke);
    pe = wave_sum(pe);
    __shared__ double s[2][4];
    if ((threadIdx.x & 63) == 0)
        {
        s[0][threadIdx.x >> 6] = ke;
        s[1][threadIdx.x >> 6] = pe;
        }
    __syncthreads();
    if (threadIdx.x == 0)
        {
        double a = 0, b = 0;
        for (int w = 0; w < int(blockDim.x >> 6); ++w)
            {
            a += s[0][w];
            b += s[1][w];
            }
        part[2 * blockIdx.x] = a;
        part[2 * blockIdx.x + 1] = b;
        }
    }

// ---------------------------------------------------------------------------
// Batched MD: one launch advances many independent systems.  One workgroup per item
// (system, chunk of 256 particles); each thread runs the single-system per-particle code.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) kb_vv_first(const md::MDSys* __restrict__ ds, const md::MDItem* __restrict__ items, double dt,
                                                   const unsigned char* __restrict__ hold)
    {
    const md::MDItem w = items[blockIdx.x];
    if (hold != nullptr && hold[w.sys])
        return;  // a held system is not integrated this step (its forces are being computed)
    const md::MDSys& d = ds[w.sys];
    const unsigned int i = w.chunk * 256u + threadIdx.x;
    if (i < d.N)
        vv_first_one(i, reinterpret_cast<double4*>(d.postype), reinterpret_cast<double4*>(d.vel),
                     reinterpret_cast<const double4*>(d.force), dt, d.L[0], d.L[1], d.L[2]);
    }

__global__ void __launch_bounds__(256) kb_vv_second(const md::MDSys* __restrict__ ds, const md::MDItem* __restrict__ items, double dt,
                                                    double gamma, double kT, uint64_t step, const unsigned char* __restrict__ hold)
    {
    const md::MDItem w = items[blockIdx.x];
    if (hold != nullptr && hold[w.sys])
        return;
    const md::MDSys& d = ds[w.sys];
    const unsigned int i = w.chunk * 256u + threadIdx.x;
    if (i < d.N)
        vv_second_one(i, reinterpret_cast<double4*>(d.vel), reinterpret_cast<const double4*>(d.force), dt, gamma, kT, d.seed, step);
    }

__global__ void __launch_bounds__(256) kb_wca(const md::MDSys* __restrict__ ds, const md::MDItem* __restrict__ items, double sigma,
                                              double epsilon)
    {
    const md::MDItem w = items[blockIdx.x];
    const md::MDSys& d = ds[w.sys];
    if (d.nlist == nullptr)
        return;
    // the inner list when it covers the WCA cutoff 2^(1/6) sigma (it holds every pair the full
    // list holds within r_inner + buffer, so it stays valid as long as the full list does)
    const bool use_inner = d.inner != nullptr && 1.122462048309373 * sigma <= d.r_inner;
    const unsigned int* n_neigh = use_inner ? d.n_inner : d.n_neigh;
    const unsigned int* nlist = use_inner ? d.inner : d.nlist;
    const uint64_t* head = use_inner ? d.head_inner : d.head;
    if (use_inner)
        {
        // a few entries per particle: one lane each, one pass (the lane groups below are for
        // the ~100-entry ESP list and would leave most lanes idle through eight passes)
        const unsigned int i = w.chunk * 256u + threadIdx.x;
        wca_group<1>(i, 0, i < d.N, reinterpret_cast<const double4*>(d.postype), d.L[0], d.L[1], d.L[2], n_neigh, nlist, head, sigma,
                     epsilon, reinterpret_cast<double4*>(d.force), d.virial, d.pitch);
        return;
        }
    // the item's 256 particles in WCA_LANES passes of 256 / WCA_LANES particle groups
    constexpr unsigned int PER_PASS = 256u / unsigned(WCA_LANES);
    for (unsigned int pass = 0; pass < unsigned(WCA_LANES); ++pass)
        {
        const unsigned int i = w.chunk * 256u + pass * PER_PASS + threadIdx.x / unsigned(WCA_LANES);
        wca_group<WCA_LANES>(i, int(threadIdx.x % WCA_LANES), i < d.N, reinterpret_cast<const double4*>(d.postype), d.L[0], d.L[1],
                             d.L[2], n_neigh, nlist, head, sigma, epsilon, reinterpret_cast<double4*>(d.force), d.virial, d.pitch);
        }
    }

//! Items (chunks of 256 particles, grouped by system) per workgroup of the batched displacement
//! checks.  A workgroup keeps one running maximum per system and sends it to the system's word
//! when the system changes or its items end, so a system of 200,000 particles meets its word
//! with ~100 atomics instead of ~780: same-address atomics serialise (the served stream's
//! predicted check took 352 us per tick with one item per workgroup).
constexpr unsigned int DISP_ITEMS = 8;

//! The batched displacement checks: for items [blockIdx.x DISP_ITEMS, + DISP_ITEMS), the
//! largest disp2(d, i) per system into out[sys] (bits of a non-negative double).  The system
//! of an item is the same for every thread, so the flushes (and their barriers) are uniform.
template<typename F>
__device__ __forceinline__ void batch_disp_max(const md::MDSys* __restrict__ ds, const md::MDItem* __restrict__ items, unsigned int n_items,
                                               unsigned long long* __restrict__ out, F disp2)
    {
    const unsigned int it0 = blockIdx.x * DISP_ITEMS;
    const unsigned int it1 = min(n_items, it0 + DISP_ITEMS);
    unsigned int cur = items[it0].sys;
    double m = 0.0;
    for (unsigned int it = it0; it < it1; ++it)
        {
        const md::MDItem w = items[it];
        if (w.sys != cur)
            {
            block_max_to(m, out + cur);
            __syncthreads();  // block_max_to's wave slots are written again by the next flush
            m = 0.0;
            cur = w.sys;
            }
        const md::MDSys& d = ds[w.sys];
        const unsigned int i = w.chunk * 256u + threadIdx.x;
        if (i < d.N && d.last_pos != nullptr)
            m = fmax(m, disp2(d, i));
        }
    block_max_to(m, out + cur);
    }

//! Per-system largest squared displacement since its last neighbour-list build (bits of a
//! non-negative double in out[sys], which the caller zeroes).
__global__ void __launch_bounds__(256) kb_max_disp(const md::MDSys* __restrict__ ds, const md::MDItem* __restrict__ items, unsigned int n_items,
                                                   unsigned long long* __restrict__ out)
    {
    batch_disp_max(ds, items, n_items, out, [](const md::MDSys& d, unsigned int i) {
        return displacement2(reinterpret_cast<const double4*>(d.postype)[i], reinterpret_cast<const double4*>(d.last_pos)[i], d.L);
    });
    }

//! Predicted displacement check: per system, the largest squared distance between the position
//! the next first half step (kb_vv_first at the same dt) will produce -- computed by the same
//! arithmetic -- and the position at the last list build.  Run at the end of a step, it tells
//! the host which lists the next step must rebuild after its drift, with no wait inside that
//! step; the decision is the one kb_max_disp would give after the drift.
__global__ void __launch_bounds__(256) kb_predict_disp(const md::MDSys* __restrict__ ds, const md::MDItem* __restrict__ items,
                                                       unsigned int n_items, double dt, unsigned long long* __restrict__ out)
    {
    batch_disp_max(ds, items, n_items, out, [dt](const md::MDSys& d, unsigned int i) {
        double4 r = reinterpret_cast<const double4*>(d.postype)[i];
        double4 v = reinterpret_cast<const double4*>(d.vel)[i];
        vv_first_values(r, v, reinterpret_cast<const double4*>(d.force)[i], dt, d.L[0], d.L[1], d.L[2]);
        return displacement2(r, reinterpret_cast<const double4*>(d.last_pos)[i], d.L);
    });
    }

//! dst row i = src row order[i], rows of `words` 4-byte words (a particle permutation); rows
//! whose index is out of range are left as they are.
__global__ void __launch_bounds__(256) k_gather_rows(unsigned int n, unsigned int words, const unsigned int* __restrict__ order,
                                                     const unsigned int* __restrict__ src, unsigned int* __restrict__ dst)
    {
    const size_t total = size_t(n) * words;
    for (size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < total; idx += size_t(gridDim.x) * blockDim.x)
        {
        const unsigned int r = unsigned(idx / words);
        const unsigned int w = unsigned(idx - size_t(r) * words);
        const unsigned int j = order[r];
        if (j < n)
            dst[idx] = src[size_t(j) * words + w];
        }
    }

    } // namespace dev

namespace md
{
using namespace esp::dev;

static unsigned int nblocks(unsigned int n)
    {
    return (n + 255u) / 256u;
    }

hipError_t cell_count(unsigned int N, const double* pos, const CellGeom& cg, unsigned int* count, unsigned int* cell, unsigned int* slot,
                      hipStream_t s)
    {
    if (N == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_cell_count, dim3(nblocks(N)), dim3(256), 0, s, N, reinterpret_cast<const double4*>(pos), cg, count, cell, slot);
    return hipGetLastError();
    }

hipError_t cell_fill(unsigned int N, const double* pos, const CellGeom& cg, const unsigned int* cell, const unsigned int* slot,
                     const unsigned int* start, unsigned int* unsorted, unsigned int* members, double* spos, hipStream_t s)
    {
    if (N == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_cell_fill, dim3(nblocks(N)), dim3(256), 0, s, N, cell, slot, start, unsorted);
    hipLaunchKernelGGL(k_cell_order, dim3(nblocks(N)), dim3(256), 0, s, N, reinterpret_cast<const double4*>(pos), cg, cell, start,
                       unsorted, members, reinterpret_cast<double4*>(spos));
    return hipGetLastError();
    }

hipError_t build_nlist(unsigned int N, const double* spos, const CellGeom& cg, const unsigned int* cell, const unsigned int* start,
                       const unsigned int* members, double r_list, const unsigned int* n_ex, const unsigned int* ex_list,
                       unsigned int ex_pitch, unsigned int nmax, unsigned int* n_neigh, unsigned int* nlist, uint64_t* head,
                       unsigned int* overflow, const InnerList& inner, hipStream_t s)
    {
    if (N == 0)
        return hipSuccess;
    if (inner.r2 > 0.0 && (!inner.n || !inner.list || !inner.head || inner.nmax == 0))
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_nlist, dim3(nblocks(N * unsigned(dev::NLIST_LANES))), dim3(256), 0, s, N, reinterpret_cast<const double4*>(spos),
                       cg, cell, start, members, r_list * r_list, n_ex, ex_list, ex_pitch, nmax, n_neigh, nlist, head, overflow, inner);
    return hipGetLastError();
    }

hipError_t wca_forces(unsigned int N, const double* pos, double Lx, double Ly, double Lz, const unsigned int* n_neigh,
                      const unsigned int* nlist, const uint64_t* head, double sigma, double epsilon, double* force, double* virial,
                      size_t pitch, hipStream_t s)
    {
    if (N == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_wca, dim3(nblocks(N)), dim3(256), 0, s, N, reinterpret_cast<const double4*>(pos), Lx, Ly, Lz, n_neigh, nlist, head,
                       sigma, epsilon, reinterpret_cast<double4*>(force), virial, pitch);
    return hipGetLastError();
    }

hipError_t vv_first(unsigned int N, double* pos, double* vel, const double* force, double dt, double Lx, double Ly, double Lz, hipStream_t s)
    {
    if (N == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_vv_first, dim3(nblocks(N)), dim3(256), 0, s, N, reinterpret_cast<double4*>(pos), reinterpret_cast<double4*>(vel),
                       reinterpret_cast<const double4*>(force), dt, Lx, Ly, Lz);
    return hipGetLastError();
    }

hipError_t vv_second(unsigned int N, double* vel, const double* force, double dt, double gamma, double kT, uint64_t seed, uint64_t step,
                     hipStream_t s)
    {
    if (N == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_vv_second, dim3(nblocks(N)), dim3(256), 0, s, N, reinterpret_cast<double4*>(vel),
                       reinterpret_cast<const double4*>(force), dt, gamma, kT, seed, step);
    return hipGetLastError();
    }

hipError_t max_displacement(unsigned int N, const double* pos, const double* last, double Lx, double Ly, double Lz, unsigned long long* out,
                            hipStream_t s)
    {
    if (N == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_max_disp, dim3(std::min(nblocks(N), dev::MAX_DISP_BLOCKS)), dim3(256), 0, s, N, reinterpret_cast<const double4*>(pos),
                       reinterpret_cast<const double4*>(last), Lx, Ly, Lz, out);
    return hipGetLastError();
    }

hipError_t copy4(unsigned int N, const double* src, double* dst, hipStream_t s)
    {
    if (N == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_copy4, dim3(nblocks(N)), dim3(256), 0, s, N, reinterpret_cast<const double4*>(src), reinterpret_cast<double4*>(dst));
    return hipGetLastError();
    }

hipError_t energies(unsigned int N, const double* vel, const double* force, double* part, hipStream_t s)
    {
    if (N == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_energies, dim3(nblocks(N)), dim3(256), 0, s, N, reinterpret_cast<const double4*>(vel),
                       reinterpret_cast<const double4*>(force), part);
    return hipGetLastError();
    }

hipError_t batch_vv_first(const MDSys* d, const MDItem* items, unsigned int n_items, double dt, const unsigned char* hold,
                          hipStream_t s)
    {
    if (n_items == 0)
        return hipSuccess;
    hipLaunchKernelGGL(kb_vv_first, dim3(n_items), dim3(256), 0, s, d, items, dt, hold);
    return hipGetLastError();
    }

hipError_t batch_vv_second(const MDSys* d, const MDItem* items, unsigned int n_items, double dt, double gamma, double kT, uint64_t step,
                           const unsigned char* hold, hipStream_t s)
    {
    if (n_items == 0)
        return hipSuccess;
    hipLaunchKernelGGL(kb_vv_second, dim3(n_items), dim3(256), 0, s, d, items, dt, gamma, kT, step, hold);
    return hipGetLastError();
    }

hipError_t batch_wca(const MDSys* d, const MDItem* items, unsigned int n_items, double sigma, double epsilon, hipStream_t s)
    {
    if (n_items == 0)
        return hipSuccess;
    hipLaunchKernelGGL(kb_wca, dim3(n_items), dim3(256), 0, s, d, items, sigma, epsilon);
    return hipGetLastError();
    }

hipError_t batch_max_disp(const MDSys* d, const MDItem* items, unsigned int n_items, unsigned long long* out, hipStream_t s)
    {
    if (n_items == 0)
        return hipSuccess;
    hipLaunchKernelGGL(kb_max_disp, dim3((n_items + DISP_ITEMS - 1) / DISP_ITEMS), dim3(256), 0, s, d, items, n_items, out);
    return hipGetLastError();
    }

hipError_t gather_rows(unsigned int n, unsigned int words, const unsigned int* order, const void* src, void* dst, hipStream_t s)
    {
    const size_t total = size_t(n) * words;
    if (total == 0)
        return hipSuccess;
    const unsigned int nb = unsigned(std::min<size_t>((total + 255) / 256, 8192));
    hipLaunchKernelGGL(k_gather_rows, dim3(nb), dim3(256), 0, s, n, words, order, static_cast<const unsigned int*>(src),
                       static_cast<unsigned int*>(dst));
    return hipGetLastError();
    }

hipError_t batch_predict_disp(const MDSys* d, const MDItem* items, unsigned int n_items, double dt, unsigned long long* out,
                              hipStream_t s)
    {
    if (n_items == 0)
        return hipSuccess;
    hipLaunchKernelGGL(kb_predict_disp, dim3((n_items + DISP_ITEMS - 1) / DISP_ITEMS), dim3(256), 0, s, d, items, n_items, dt, out);
    return hipGetLastError();
    }

    } // namespace md
    } // namespace esp
