// Test probe: include/EvaluatorPairPSWF.h compiled by the host compiler (the CPU path of
// PotentialPair<EvaluatorPairPSWF>) with its pybind11 dict constructor, exposed to pytest.
// Built by tests/test_evaluator.py with g++.
#define ESP_EVALUATOR_PYBIND11
#include "EvaluatorPairPSWF.h"

#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <cstdint>

namespace py = pybind11;
using Eval = esp_hoomd::EvaluatorPairPSWF;

// evalForceAndEnergy over arrays of rsq and qiqj; returns (ok, force_divr, pair_eng)
static py::tuple eval_pairs(py::dict params, double rcutsq, bool shift, py::array_t<double> rsq, py::array_t<double> qiqj)
    {
    const Eval::param_type p(params);
    const auto r = rsq.unchecked<1>();
    const auto q = qiqj.unchecked<1>();
    const py::ssize_t n = r.shape(0);
    py::array_t<int> ok(n);
    py::array_t<double> f(n), e(n);
    auto okm = ok.mutable_unchecked<1>();
    auto fm = f.mutable_unchecked<1>();
    auto em = e.mutable_unchecked<1>();
    for (py::ssize_t i = 0; i < n; ++i)
        {
        Eval ev(r(i), rcutsq, p);
        if (Eval::needsCharge())
            ev.setCharge(1.0, q(i));
        double fd = -7.0, pe = -7.0;  // left untouched when the pair is rejected
        okm(i) = ev.evalForceAndEnergy(fd, pe, shift) ? 1 : 0;
        fm(i) = fd;
        em(i) = pe;
        }
    return py::make_tuple(ok, f, e);
    }

static py::dict round_trip(py::dict params)
    {
    return Eval::param_type(params).asDict();
    }

static py::dict facts(py::dict params)
    {
    const Eval::param_type p(params);
    Eval ev(1.0, 4.0, p);
    py::dict d;
    d["name"] = Eval::getName();
    d["needs_charge"] = Eval::needsCharge();
    d["needs_diameter"] = Eval::needsDiameter();
    d["useless"] = ev.areInteractionParametersUseless();
    d["lrc_energy"] = ev.evalEnergyLRCIntegral();
    d["lrc_pressure"] = ev.evalPressureLRCIntegral();
    d["param_size"] = sizeof(Eval::param_type);
    d["param_align"] = alignof(Eval::param_type);
    bool threw = false;
    try
        {
        (void)ev.getShapeSpec();
        }
    catch (const std::runtime_error&)
        {
        threw = true;
        }
    d["shape_spec_throws"] = threw;
    return d;
    }

// the table address as the dict carries it
static std::uintptr_t address_of(py::array_t<double> a)
    {
    return reinterpret_cast<std::uintptr_t>(a.data());
    }

PYBIND11_MODULE(evaluator_probe, m)
    {
    m.def("eval_pairs", &eval_pairs);
    m.def("round_trip", &round_trip);
    m.def("facts", &facts);
    m.def("address_of", &address_of);
    }
