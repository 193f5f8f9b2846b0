// ESPForceComputeMI355X -- the HOOMD-blue force compute of the ESP long-range part, backed by the
// MI355X library (libesp_mi355x.so, include/esp_mi355x.h).  It keeps the class surface of the
// reference's ESPForceCompute / ESPForceComputeGPU (ESPForceCompute.h:90-233,
// ESPForceComputeGPU.cc:457-481): the (sysdef, nlist, group) constructor, setParams(nx, ny, nz,
// order, kappa, rcut, alpha, n_table), the getters, the introspection table and invalidate(),
// and computes through esp_compute_forces on the device arrays HOOMD owns.
//
// Built against HOOMD-blue (v4/v5 headers, HIP build) by CMakeLists.txt in this directory;
// HOOMD is not part of this repository.
#pragma once

#include <hoomd/ForceCompute.h>
#include <hoomd/GlobalArray.h>
#include <hoomd/ParticleGroup.h>
#include <hoomd/md/NeighborList.h>

#include <pybind11/pybind11.h>

#include <memory>
#include <vector>

#include "esp_mi355x.h"

namespace esp_hoomd
{

class ESPForceComputeMI355X : public hoomd::ForceCompute
    {
    public:
    //! ESPForceCompute.cc:53-69
    ESPForceComputeMI355X(std::shared_ptr<hoomd::SystemDefinition> sysdef, std::shared_ptr<hoomd::md::NeighborList> nlist,
                          std::shared_ptr<hoomd::ParticleGroup> group);
    ~ESPForceComputeMI355X() override;

    //! ESPForceCompute::setParams (ESPForceCompute.cc:71-89); precision is ESP_PREC_FP64 or
    //! ESP_PREC_MIXED (an addition: the reference has FP64 only)
    void setParams(unsigned int nx, unsigned int ny, unsigned int nz, unsigned int order, hoomd::Scalar kappa,
                   hoomd::Scalar rcut, hoomd::Scalar alpha, unsigned int n_table, int precision);

    // getters (ESPForceCompute.h:108-118)
    pybind11::tuple getResolution();
    unsigned int getOrder();
    hoomd::Scalar getKappa();
    hoomd::Scalar getRCut();
    hoomd::Scalar getAlpha();
    hoomd::Scalar getQSum();
    hoomd::Scalar getQ2Sum();
    unsigned int getTableSize();
    //! host address of the split table (ESPForceCompute.cc:129-132), for the pair evaluator
    uintptr_t getTablePtr();
    //! device address of the same table, for the pair evaluator on the GPU
    uintptr_t getTableDevicePtr();
    //! segments of the split table
    unsigned int getTableSegments();
    //! ESPForceCompute::getTableEntry (ESPForceCompute.cc:336-372); throws std::out_of_range
    esp_table_entry getTableEntry(unsigned int i);
    //! ESPForceCompute::invalidate (ESPForceCompute.h:120)
    void invalidate();
    //! fused (default): this compute also evaluates the pair correction and the exclusion
    //! correction over the neighbour list (and registers r_cut with it), so no PotentialPairPSWF
    //! is attached; not fused: mesh part only, the pair part comes from PotentialPairPSWF[GPU]
    //! (systems without exclusions)
    void setFused(bool fused);

    protected:
    //! ESPForceCompute::computeForces (ESPForceCompute.cc:96-127)
    void computeForces(uint64_t timestep) override;

    private:
    void check(int rc) const;

    std::shared_ptr<hoomd::md::NeighborList> m_nlist;
    std::shared_ptr<hoomd::ParticleGroup> m_group;
    esp_ctx* m_ctx = nullptr;
    double* m_ev_dev = nullptr;  //!< device double[7]: external energy and virial
    esp_params m_params {};
    bool m_fused = true;
    std::shared_ptr<hoomd::GlobalArray<hoomd::Scalar>> m_r_cut_nlist;  //!< r_cut per type pair for the list
    void publish_r_cut();
    };

//! Export ESPForceComputeMI355X and ESPTableEntry to Python (ESPForceCompute.cc:447-508).
void export_ESPForceComputeMI355X(pybind11::module& m);

} // namespace esp_hoomd
