// HOOMD's generic GPU pair kernel instantiated for EvaluatorPairPSWF: the driver behind
// PotentialPairPSWFGPU (module.cc:79 of the reference, where it was exported but never
// instantiated).  Compiled by hipcc for gfx950 against HOOMD-blue's headers.
#include <hoomd/md/PotentialPairGPU.cuh>

#include "EvaluatorPairPSWF.h"

namespace hoomd
{
namespace md
{
namespace kernel
{
template hipError_t gpu_compute_pair_forces<esp_hoomd::EvaluatorPairPSWF>(const pair_args_t& pair_args,
                                                                         const esp_hoomd::EvaluatorPairPSWF::param_type* d_params);
} // namespace kernel
} // namespace md
} // namespace hoomd
