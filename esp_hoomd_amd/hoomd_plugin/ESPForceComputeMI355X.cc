// ESPForceComputeMI355X: HOOMD-blue arrays handed to the MI355X library through its C ABI.
#include "ESPForceComputeMI355X.h"

#include <hoomd/GPUArray.h>

#include <sstream>
#include <stdexcept>
#include <string>

namespace esp_hoomd
{

ESPForceComputeMI355X::ESPForceComputeMI355X(std::shared_ptr<hoomd::SystemDefinition> sysdef,
                                             std::shared_ptr<hoomd::md::NeighborList> nlist,
                                             std::shared_ptr<hoomd::ParticleGroup> group)
    : hoomd::ForceCompute(sysdef), m_nlist(nlist), m_group(group)
    {
    if (!m_exec_conf->isCUDAEnabled())
        throw std::runtime_error("ESPForceComputeMI355X: the MI355X ESP library needs a GPU execution configuration");
    check(esp_device_malloc(reinterpret_cast<void**>(&m_ev_dev), 7 * sizeof(double)));
    // the reference's defaults until setParams (ESPForceCompute.cc:53-69)
    m_params.nx = m_params.ny = m_params.nz = 32;
    m_params.order = 6;
    m_params.kappa = 1.0;
    m_params.rcut = 1.0;
    m_params.n_table = 4096;
    m_params.precision = ESP_PREC_FP64;
    m_params.split = ESP_SPLIT_PSWF;
    // the neighbour list must cover r_cut for the fused pair correction, as PotentialPair's does
    const unsigned int nt = m_pdata->getNTypes();
    m_r_cut_nlist = std::make_shared<hoomd::GlobalArray<hoomd::Scalar>>(nt * nt, m_exec_conf);
    m_nlist->addRCutMatrix(m_r_cut_nlist);
    }

void ESPForceComputeMI355X::publish_r_cut()
    {
    hoomd::ArrayHandle<hoomd::Scalar> rc(*m_r_cut_nlist, hoomd::access_location::host, hoomd::access_mode::overwrite);
    for (size_t k = 0; k < m_r_cut_nlist->getNumElements(); ++k)
        rc.data[k] = m_fused ? m_params.rcut : hoomd::Scalar(0);
    m_nlist->notifyRCutMatrixChange();
    }

void ESPForceComputeMI355X::setFused(bool fused)
    {
    m_fused = fused;
    publish_r_cut();
    }

ESPForceComputeMI355X::~ESPForceComputeMI355X()
    {
    m_nlist->removeRCutMatrix(m_r_cut_nlist);
    if (m_ctx)
        esp_destroy(m_ctx);
    if (m_ev_dev)
        esp_device_free(m_ev_dev);
    }

void ESPForceComputeMI355X::check(int rc) const
    {
    if (rc == ESP_OK)
        return;
    const std::string msg = std::string("ESP: ") + esp_last_error(m_ctx);
    if (rc == ESP_ERR_RANGE)
        throw std::out_of_range(msg);
    throw std::runtime_error(msg);
    }

void ESPForceComputeMI355X::setParams(unsigned int nx, unsigned int ny, unsigned int nz, unsigned int order,
                                      hoomd::Scalar kappa, hoomd::Scalar rcut, hoomd::Scalar alpha, unsigned int n_table,
                                      int precision)
    {
    esp_params p = m_params;
    p.nx = nx;
    p.ny = ny;
    p.nz = nz;
    p.order = order;
    p.kappa = kappa;
    p.rcut = rcut;
    p.alpha = alpha;
    p.n_table = n_table;
    p.precision = precision;
    check(m_ctx ? esp_set_params(m_ctx, &p) : esp_create(&m_ctx, &p));
    m_params = p;
    publish_r_cut();
    }

void ESPForceComputeMI355X::computeForces(uint64_t timestep)
    {
    if (!m_ctx)
        throw std::runtime_error("ESP: setParams must be called before the first step");
    m_nlist->compute(timestep);
    using hoomd::access_location;
    using hoomd::access_mode;
    hoomd::ArrayHandle<hoomd::Scalar4> pos(m_pdata->getPositions(), access_location::device, access_mode::read);
    hoomd::ArrayHandle<hoomd::Scalar> q(m_pdata->getCharges(), access_location::device, access_mode::read);
    hoomd::ArrayHandle<unsigned int> nn(m_nlist->getNNeighArray(), access_location::device, access_mode::read);
    hoomd::ArrayHandle<unsigned int> nl(m_nlist->getNListArray(), access_location::device, access_mode::read);
    hoomd::ArrayHandle<size_t> head(m_nlist->getHeadList(), access_location::device, access_mode::read);
    hoomd::ArrayHandle<unsigned int> nex(m_nlist->getNExArray(), access_location::device, access_mode::read);
    hoomd::ArrayHandle<unsigned int> exl(m_nlist->getExListArray(), access_location::device, access_mode::read);
    hoomd::ArrayHandle<unsigned int> grp(m_group->getIndexArray(), access_location::device, access_mode::read);
    hoomd::ArrayHandle<hoomd::Scalar4> f(m_force, access_location::device, access_mode::overwrite);
    hoomd::ArrayHandle<hoomd::Scalar> v(m_virial, access_location::device, access_mode::overwrite);
    const hoomd::BoxDim box = m_pdata->getGlobalBox();
    const hoomd::Scalar3 L = box.getL();

    esp_system s {};
    s.N = m_pdata->getN();
    s.postype = reinterpret_cast<const double*>(pos.data);
    s.charge = q.data;
    s.group_size = m_group->getNumMembers();
    s.group_idx = grp.data;
    s.box[0] = L.x;
    s.box[1] = L.y;
    s.box[2] = L.z;
    s.box[3] = box.getTiltFactorXY();
    s.box[4] = box.getTiltFactorXZ();
    s.box[5] = box.getTiltFactorYZ();
    s.n_neigh = nn.data;
    s.nlist = nl.data;
    s.head_list = reinterpret_cast<const uint64_t*>(head.data);
    s.n_ex = nex.data;
    s.ex_list = exl.data;
    s.ex_pitch = m_nlist->getExListIndexer().getW();

    esp_out o {};
    o.force = reinterpret_cast<double*>(f.data);
    o.virial = v.data;
    o.virial_pitch = m_virial.getPitch();
    o.ext_energy_virial = m_ev_dev;
    o.flags = m_fused ? ESP_COMPUTE_ALL : (ESP_COMPUTE_MESH | ESP_COMPUTE_ENERGY_VIRIAL);
    // HOOMD runs its kernels on the default stream
    check(esp_compute_forces(m_ctx, timestep, &s, &o, nullptr));

    // the mesh, self and background energy and virial are not per particle: ForceCompute's
    // external energy and virial (the reference zeroes them, ESPForceCompute.cc:124-126)
    double ev[7];
    check(esp_memcpy_dtoh(ev, m_ev_dev, sizeof(ev)));
    m_external_energy = ev[0];
    for (int k = 0; k < 6; ++k)
        m_external_virial[k] = ev[1 + k];
    }

pybind11::tuple ESPForceComputeMI355X::getResolution()
    {
    return pybind11::make_tuple(m_params.nx, m_params.ny, m_params.nz);
    }
unsigned int ESPForceComputeMI355X::getOrder()
    {
    return m_params.order;
    }
hoomd::Scalar ESPForceComputeMI355X::getKappa()
    {
    return m_params.kappa;
    }
hoomd::Scalar ESPForceComputeMI355X::getRCut()
    {
    return m_params.rcut;
    }
hoomd::Scalar ESPForceComputeMI355X::getAlpha()
    {
    return m_params.alpha;
    }
hoomd::Scalar ESPForceComputeMI355X::getQSum()
    {
    double qs = 0, q2 = 0;
    check(esp_get_charge_sums(m_ctx, &qs, &q2));
    return qs;
    }
hoomd::Scalar ESPForceComputeMI355X::getQ2Sum()
    {
    double qs = 0, q2 = 0;
    check(esp_get_charge_sums(m_ctx, &qs, &q2));
    return q2;
    }
unsigned int ESPForceComputeMI355X::getTableSize()
    {
    unsigned int n = 0;
    check(esp_get_table_size(m_ctx, &n));
    return n;
    }
uintptr_t ESPForceComputeMI355X::getTablePtr()
    {
    const double *h = nullptr, *d = nullptr;
    unsigned int n = 0;
    check(esp_get_pair_table(m_ctx, &h, &d, &n));
    return reinterpret_cast<uintptr_t>(h);
    }
uintptr_t ESPForceComputeMI355X::getTableDevicePtr()
    {
    const double *h = nullptr, *d = nullptr;
    unsigned int n = 0;
    check(esp_get_pair_table(m_ctx, &h, &d, &n));
    return reinterpret_cast<uintptr_t>(d);
    }
unsigned int ESPForceComputeMI355X::getTableSegments()
    {
    const double *h = nullptr, *d = nullptr;
    unsigned int n = 0;
    check(esp_get_pair_table(m_ctx, &h, &d, &n));
    return n;
    }
esp_table_entry ESPForceComputeMI355X::getTableEntry(unsigned int i)
    {
    esp_table_entry e {};
    check(esp_get_table_entry(m_ctx, i, &e));
    return e;
    }
void ESPForceComputeMI355X::invalidate()
    {
    if (m_ctx)
        check(esp_invalidate(m_ctx));
    }

void export_ESPForceComputeMI355X(pybind11::module& m)
    {
    pybind11::class_<esp_table_entry>(m, "ESPTableEntry")
        .def_readonly("r_lo", &esp_table_entry::r_lo)
        .def_readonly("dr_inv", &esp_table_entry::dr_inv)
        .def_readonly("potential_coeff4", &esp_table_entry::potential_coeff4)
        .def_readonly("force_coeff4", &esp_table_entry::force_coeff4)
        .def_property_readonly("potential_coeffs",
                               [](const esp_table_entry& e)
                               {
                                   return pybind11::make_tuple(e.potential_coeffs[0], e.potential_coeffs[1], e.potential_coeffs[2],
                                                               e.potential_coeffs[3]);
                               })
        .def_property_readonly("force_coeffs",
                               [](const esp_table_entry& e)
                               { return pybind11::make_tuple(e.force_coeffs[0], e.force_coeffs[1], e.force_coeffs[2], e.force_coeffs[3]); })
        .def("__repr__",
             [](const esp_table_entry& e)
             {
                 std::ostringstream s;
                 s << "<ESPTableEntry r_lo=" << e.r_lo << " dr_inv=" << e.dr_inv << ">";
                 return s.str();
             });

    pybind11::class_<ESPForceComputeMI355X, hoomd::ForceCompute, std::shared_ptr<ESPForceComputeMI355X>>(m, "ESPForceComputeMI355X")
        .def(pybind11::init<std::shared_ptr<hoomd::SystemDefinition>, std::shared_ptr<hoomd::md::NeighborList>,
                            std::shared_ptr<hoomd::ParticleGroup>>(),
             pybind11::arg("sysdef"), pybind11::arg("nlist"), pybind11::arg("group"))
        .def("setParams", &ESPForceComputeMI355X::setParams, pybind11::arg("nx"), pybind11::arg("ny"), pybind11::arg("nz"),
             pybind11::arg("order"), pybind11::arg("kappa"), pybind11::arg("rcut"), pybind11::arg("alpha") = 0.0,
             pybind11::arg("n_table") = 4096u, pybind11::arg("precision") = int(ESP_PREC_FP64))
        .def("getResolution", &ESPForceComputeMI355X::getResolution)
        .def("getOrder", &ESPForceComputeMI355X::getOrder)
        .def("getKappa", &ESPForceComputeMI355X::getKappa)
        .def("getRCut", &ESPForceComputeMI355X::getRCut)
        .def("getAlpha", &ESPForceComputeMI355X::getAlpha)
        .def("getQSum", &ESPForceComputeMI355X::getQSum)
        .def("getQ2Sum", &ESPForceComputeMI355X::getQ2Sum)
        .def("getTableSize", &ESPForceComputeMI355X::getTableSize)
        .def("getTablePtr", &ESPForceComputeMI355X::getTablePtr)
        .def("getTableDevicePtr", &ESPForceComputeMI355X::getTableDevicePtr)
        .def("getTableSegments", &ESPForceComputeMI355X::getTableSegments)
        .def("getTableEntry", &ESPForceComputeMI355X::getTableEntry, pybind11::arg("index"))
        .def("invalidate", &ESPForceComputeMI355X::invalidate)
        .def("setFused", &ESPForceComputeMI355X::setFused);
    }

} // namespace esp_hoomd
