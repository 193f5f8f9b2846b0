// Short-range ESP correction over a HOOMD-layout full neighbour list, with the excluded-pair
// correction fused in.  Each particle is handled by a group of G lanes; the ten per-particle
// sums (force, energy, virial) are reduced with cross-lane shuffles inside the wave.
//
// Reference counterparts: EvaluatorPairPSWF::evalForceAndEnergy (EvaluatorPairPSWF.h:274-295)
// called by HOOMD's PotentialPairGPU, and gpu_esp_fix_exclusions_kernel
// (ESPForceComputeGPU.cu:391-471).
#include "esp_device.h"
#include "esp_kernels.h"

namespace esp
{
namespace dev
{

__device__ __forceinline__ double min_image(double d, double L, double invL)
    {
    return d - L * rint(d * invL);
    }

template<typename R, int G>
__global__ void __launch_bounds__(256) k_pair(kernel::PairArgs a)
    {
    const unsigned int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned int i = gid / G;
    const int sub = int(gid % G);
    const bool active = i < a.N;
    const R* __restrict__ tab = reinterpret_cast<const R*>(a.split_tab);
    const double invLx = 1.0 / a.Lx, invLy = 1.0 / a.Ly, invLz = 1.0 / a.Lz;
    const R rc = R(a.rc);
    const R inv_rc = R(1.0 / a.rc);
    const R rc2 = rc * rc;

    R f[3] = {R(0), R(0), R(0)};
    R e = R(0);
    R v[6] = {R(0), R(0), R(0), R(0), R(0), R(0)};
    if (active)
        {
        const double4 pi = reinterpret_cast<const double4*>(a.postype)[i];
        const double qi = a.charge[i];
        if (a.nlist != nullptr && qi != 0.0 && a.rc >= 1.0e-12)
            {
            const unsigned int nn = a.n_neigh[i];
            const uint64_t head = a.head_list[i];
            for (unsigned int k = sub; k < nn; k += G)
                {
                const unsigned int j = a.nlist[head + k];
                const double4 pj = reinterpret_cast<const double4*>(a.postype)[j];
                const R qiqj = R(qi * a.charge[j]);
                const R dx = R(min_image(pi.x - pj.x, a.Lx, invLx));
                const R dy = R(min_image(pi.y - pj.y, a.Ly, invLy));
                const R dz = R(min_image(pi.z - pj.z, a.Lz, invLz));
                const R rsq = dx * dx + dy * dy + dz * dz;
                if (rsq >= rc2 || qiqj == R(0))
                    continue;
                const R r = sqrt(rsq);
                R Lr, nd;
                split_lookup<R>(tab, a.n_segs, r, rc, inv_rc, Lr, nd);
                const R fdivr = qiqj * nd / r;
                f[0] += fdivr * dx;
                f[1] += fdivr * dy;
                f[2] += fdivr * dz;
                e += R(0.5) * qiqj * Lr;
                const R hf = R(0.5) * fdivr;
                v[0] += hf * dx * dx;
                v[1] += hf * dx * dy;
                v[2] += hf * dx * dz;
                v[3] += hf * dy * dy;
                v[4] += hf * dy * dz;
                v[5] += hf * dz * dz;
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
                const double4 pj = reinterpret_cast<const double4*>(a.postype)[j];
                const R qiqj = R(qi * a.charge[j]);
                const R dx = R(min_image(pi.x - pj.x, a.Lx, invLx));
                const R dy = R(min_image(pi.y - pj.y, a.Ly, invLy));
                const R dz = R(min_image(pi.z - pj.z, a.Lz, invLz));
                const R rsq = dx * dx + dy * dy + dz * dz;
                if (rsq < R(1.0e-20))
                    continue;
                const R r = sqrt(rsq);
                R Lr = R(0), nd = R(0);
                if (r <= rc)
                    split_lookup<R>(tab, a.n_segs, r, rc, inv_rc, Lr, nd);
                const R inv4pir = R(0.079577471545947673) / r;
                const R dE = qiqj * (Lr - inv4pir);
                const R fdivr = qiqj * (nd - inv4pir / r) / r;
                f[0] += fdivr * dx;
                f[1] += fdivr * dy;
                f[2] += fdivr * dz;
                e += R(0.5) * dE;
                const R hf = R(0.5) * fdivr;
                v[0] += hf * dx * dx;
                v[1] += hf * dx * dy;
                v[2] += hf * dx * dz;
                v[3] += hf * dy * dy;
                v[4] += hf * dy * dz;
                v[5] += hf * dz * dz;
                }
            }
        }
    // reduce over the G lanes of this particle (all lanes of the wave take part)
#pragma unroll
    for (int c = 0; c < 3; ++c)
        f[c] = group_sum<G>(f[c]);
    e = group_sum<G>(e);
#pragma unroll
    for (int c = 0; c < 6; ++c)
        v[c] = group_sum<G>(v[c]);
    if (active && sub == 0)
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

    } // namespace dev

namespace kernel
{
template<typename R> hipError_t pair_forces(const PairArgs& a, int lanes_per_particle, hipStream_t s)
    {
    const size_t threads = size_t(a.N) * lanes_per_particle;
    const unsigned int nb = unsigned((threads + 255) / 256);
    if (nb == 0)
        return hipSuccess;
    switch (lanes_per_particle)
        {
        case 1: hipLaunchKernelGGL((dev::k_pair<R, 1>), dim3(nb), dim3(256), 0, s, a); break;
        case 2: hipLaunchKernelGGL((dev::k_pair<R, 2>), dim3(nb), dim3(256), 0, s, a); break;
        case 4: hipLaunchKernelGGL((dev::k_pair<R, 4>), dim3(nb), dim3(256), 0, s, a); break;
        case 8: hipLaunchKernelGGL((dev::k_pair<R, 8>), dim3(nb), dim3(256), 0, s, a); break;
        case 16: hipLaunchKernelGGL((dev::k_pair<R, 16>), dim3(nb), dim3(256), 0, s, a); break;
        case 32: hipLaunchKernelGGL((dev::k_pair<R, 32>), dim3(nb), dim3(256), 0, s, a); break;
        default: return hipErrorInvalidValue;
        }
    return hipGetLastError();
    }
template hipError_t pair_forces<double>(const PairArgs&, int, hipStream_t);
template hipError_t pair_forces<float>(const PairArgs&, int, hipStream_t);
    } // namespace kernel
    } // namespace esp
