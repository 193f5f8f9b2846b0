// pybind11 module _esp of the HOOMD-blue plugin (the reference's module.cc:57-84): the long-range
// force compute backed by the MI355X library and the short-range PotentialPair over
// EvaluatorPairPSWF, CPU and GPU.
#include <hoomd/md/PotentialPair.h>
#include <hoomd/md/PotentialPairGPU.h>

#include <pybind11/pybind11.h>

#include "ESPForceComputeMI355X.h"
#include "EvaluatorPairPSWF.h"

PYBIND11_MODULE(_esp, m)
    {
    m.doc() = "ESP (Ewald summation with prolates) on MI355X: long-range mesh force and short-range PSWF pair correction.";
    esp_hoomd::export_ESPForceComputeMI355X(m);
    hoomd::md::detail::export_PotentialPair<esp_hoomd::EvaluatorPairPSWF>(m, "PotentialPairPSWF");
    hoomd::md::detail::export_PotentialPairGPU<esp_hoomd::EvaluatorPairPSWF>(m, "PotentialPairPSWFGPU");
    m.attr("gpu_enabled") = pybind11::bool_(true);
    }
