// HOOMD's generic pair loop over the EvaluatorPairPSWF of include/EvaluatorPairPSWF.h: the path
// PotentialPairGPU<EvaluatorPairPSWF> takes in the reference plugin (module.cc:79).  One thread
// per particle walks its row of the full neighbour list (HOOMD layout), builds the evaluator
// for each pair from the squared minimum-image distance, and accumulates the force, half of
// each pair energy and half of each pair virial, as HOOMD's full-list pair kernel does.
// Outputs are overwritten.  This is the drop-in evaluator path; the library's own pair pass
// (k_pair) computes the same correction faster, fused with the exclusions.
#include "../../include/EvaluatorPairPSWF.h"
#include "../../include/esp_mi355x.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>

namespace esp
{
namespace kernel
{

using Eval = esp_hoomd::EvaluatorPairPSWF;

static_assert(sizeof(Eval::param_type) == sizeof(esp_pair_param), "esp_pair_param must mirror EvaluatorPairPSWF::param_type");
static_assert(offsetof(Eval::param_type, table_device) == offsetof(esp_pair_param, table_device), "param layout");
static_assert(offsetof(Eval::param_type, n_segs) == offsetof(esp_pair_param, n_segs), "param layout");

//! HOOMD's BoxDim::minImage for a box centred on the origin with tilt factors xy, xz, yz:
//! lattice vector a3 = (xz Lz, yz Lz, Lz) by rint(d_z / Lz), then a2 = (xy Ly, Ly, 0) by
//! rint(d_y / Ly), then a1 = (Lx, 0, 0) by rint(d_x / Lx).
__device__ __forceinline__ void min_image(double& dx, double& dy, double& dz, double Lx, double Ly, double Lz, double xy, double xz,
                                          double yz)
    {
    const double iz = rint(dz * (1.0 / Lz));
    dz -= Lz * iz;
    dy -= yz * Lz * iz;
    dx -= xz * Lz * iz;
    const double iy = rint(dy * (1.0 / Ly));
    dy -= Ly * iy;
    dx -= xy * Ly * iy;
    dx -= Lx * rint(dx * (1.0 / Lx));
    }

__global__ void __launch_bounds__(256) k_potential_pair(unsigned int N, const double4* __restrict__ postype,
                                                        const double* __restrict__ charge, const unsigned int* __restrict__ n_neigh,
                                                        const unsigned int* __restrict__ nlist, const uint64_t* __restrict__ head,
                                                        double Lx, double Ly, double Lz, double xy, double xz, double yz,
                                                        Eval::param_type param, double rcutsq,
                                                        int energy_shift, double4* __restrict__ force, double* __restrict__ virial,
                                                        size_t pitch)
    {
    const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    const double4 pi = postype[i];
    const double qi = charge[i];
    double fx = 0, fy = 0, fz = 0, e = 0;
    double v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
    const unsigned int nn = n_neigh[i];
    const uint64_t h = head[i];
    for (unsigned int k = 0; k < nn; ++k)
        {
        const unsigned int j = nlist[h + k];
        const double4 pj = postype[j];
        double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        min_image(dx, dy, dz, Lx, Ly, Lz, xy, xz, yz);
        const double rsq = dx * dx + dy * dy + dz * dz;
        Eval ev(rsq, rcutsq, param);
        if (Eval::needsCharge())
            ev.setCharge(qi, charge[j]);
        double force_divr = 0.0, pair_eng = 0.0;
        if (ev.evalForceAndEnergy(force_divr, pair_eng, energy_shift != 0))
            {
            fx += dx * force_divr;
            fy += dy * force_divr;
            fz += dz * force_divr;
            e += 0.5 * pair_eng;
            const double hv = 0.5 * force_divr;
            v0 += hv * dx * dx;
            v1 += hv * dx * dy;
            v2 += hv * dx * dz;
            v3 += hv * dy * dy;
            v4 += hv * dy * dz;
            v5 += hv * dz * dz;
            }
        }
    force[i] = make_double4(fx, fy, fz, e);
    if (virial)
        {
        virial[0 * pitch + i] = v0;
        virial[1 * pitch + i] = v1;
        virial[2 * pitch + i] = v2;
        virial[3 * pitch + i] = v3;
        virial[4 * pitch + i] = v4;
        virial[5 * pitch + i] = v5;
        }
    }

} // namespace kernel
} // namespace esp

extern "C" int esp_potential_pair_forces(const esp_system* sys, const esp_pair_param* param, double r_cut, int energy_shift,
                                         esp_out* out, esp_stream_t stream)
    {
    if (!sys || !param || !out || !out->force)
        return ESP_ERR_ARG;
    if (sys->N == 0)
        return ESP_OK;
    if (!sys->postype || !sys->charge || !sys->n_neigh || !sys->nlist || !sys->head_list)
        return ESP_ERR_ARG;
    if (!(sys->box[0] > 0 && sys->box[1] > 0 && sys->box[2] > 0) || !std::isfinite(sys->box[3]) || !std::isfinite(sys->box[4])
        || !std::isfinite(sys->box[5]))
        return ESP_ERR_ARG;
    if (out->virial && out->virial_pitch < sys->N)
        return ESP_ERR_ARG;
    if (param->n_segs == 0 || !param->table_device || !(r_cut > 0.0))
        return ESP_ERR_ARG;
    esp::kernel::Eval::param_type p;
    static_assert(sizeof(p) == sizeof(*param), "layout");
    __builtin_memcpy(&p, param, sizeof(p));
    const unsigned int blocks = (sys->N + 255u) / 256u;
    hipLaunchKernelGGL(esp::kernel::k_potential_pair, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream), sys->N,
                       reinterpret_cast<const double4*>(sys->postype), sys->charge, sys->n_neigh, sys->nlist, sys->head_list, sys->box[0],
                       sys->box[1], sys->box[2], sys->box[3], sys->box[4], sys->box[5], p, r_cut * r_cut, energy_shift, reinterpret_cast<double4*>(out->force), out->virial,
                       out->virial_pitch);
    return hipGetLastError() == hipSuccess ? ESP_OK : ESP_ERR_HIP;
    }
