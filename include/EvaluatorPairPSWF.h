/*
 * EvaluatorPairPSWF.h -- the ESP short-range pair correction as a HOOMD-blue pair evaluator.
 *
 * Replaces the reference's EvaluatorPairPSWF (EvaluatorPairPSWF.h:166-339), exported there as
 * PotentialPairPSWF / PotentialPairPSWFGPU (module.cc:73,79).  The class follows HOOMD's
 * evaluator concept, so PotentialPair<EvaluatorPairPSWF> (CPU) and HOOMD's generic GPU pair
 * kernel (hipcc, gfx950) instantiate it unchanged:
 *
 *   param_type            POD, 16-byte aligned; load_shared / allocate_shared /
 *                         set_memory_hints; dict constructor and asDict() in host builds
 *                         with pybind11 (EvaluatorPairPSWF.h:180-225)
 *   (rsq, rcutsq, param)  constructor (:231-235)
 *   needsCharge/setCharge (:238-246), needsDiameter/setDiameter (:249-254)
 *   areInteractionParametersUseless (:257-260)
 *   evalForceAndEnergy(force_divr, pair_eng, energy_shift) (:274-295): false when
 *                         rsq >= rcutsq, qiqj == 0 or rcut < 1e-12; else
 *                         pair_eng = qiqj L(r) (- qiqj V_cut when shifting),
 *                         force_divr = qiqj (-dL/dr) / r
 *   LRC integrals = 0 (:304-312), getName() == "pswf" (:319-322), getShapeSpec throws (:324-328)
 *
 * L(r) = S(r / r_c) / (4 pi r) and -dL/dr = S / (4 pi r^2) - S' / (4 pi r) come from the split
 * table of the ESP context that owns the long-range part (esp_get_pair_table in
 * esp_mi355x.h): per segment one record of 12 doubles, r_lo, dr_inv, S c0..c4, dS/ds c0..c4,
 * in the reduced distance s = r / r_c (the PSWF_Coeffs.h data, PSWF_Coeffs.h:24-9599, in one
 * interleaved array).  The lookup is the reference's (EvaluatorPairPSWF.h:113-143):
 * s clamped to [0, 1 - 1e-9], seg = min(floor(s n), n - 1), t = (s - r_lo) dr_inv, Horner
 * c4 -> c0.  Divergences from the reference, on purpose: the table is indexed by s = r / r_c
 * (the reference indexes a table built for r_c = 1 with the physical r), carries the
 * context's own split (S(0) = 1, valid for any kappa and r_c) and is passed in the
 * parameters (table_ptr / n_segs, keys the reference's dict still accepts and ignores); r is
 * an FP64 square root (the reference takes sqrtf of a double).
 *
 * The table pointer in param_type is a host address on the CPU path and a device address on
 * the GPU path (both are kept; device code reads table_device).  The table must outlive
 * every pair force that uses it.
 */
#ifndef ESP_EVALUATOR_PAIR_PSWF_H_
#define ESP_EVALUATOR_PAIR_PSWF_H_

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define ESP_EVAL_DEVICE __device__
#define ESP_EVAL_HOSTDEVICE __host__ __device__
#else
#define ESP_EVAL_DEVICE
#define ESP_EVAL_HOSTDEVICE
#endif

#if !defined(__HIPCC__)
#include <stdexcept>
#include <string>
#if defined(ESP_EVALUATOR_PYBIND11)
#include <pybind11/pybind11.h>
#endif
#endif

namespace esp_hoomd
{

//! L(r) and -dL/dr from an interleaved split table (12 doubles per segment); both 0 outside
//! (1e-12, rc].
ESP_EVAL_HOSTDEVICE inline void pswf_lookup(const double* tab, unsigned int n_segs, double r, double rc, double& L,
                                            double& negdLdr)
    {
    if (r < 1.0e-12 || r > rc || !tab || n_segs == 0)
        {
        L = 0.0;
        negdLdr = 0.0;
        return;
        }
    const double inv_rc = 1.0 / rc;
    const double su = r * inv_rc;
    const double s = su < 0.0 ? 0.0 : (su > 1.0 - 1.0e-9 ? 1.0 - 1.0e-9 : su);
    unsigned int seg = static_cast<unsigned int>(s * double(n_segs));
    seg = seg >= n_segs ? n_segs - 1u : seg;
    const double* c = tab + static_cast<size_t>(seg) * 12u;
    const double t = (su - c[0]) * c[1];
    double S = c[6];
    S = c[5] + t * S;
    S = c[4] + t * S;
    S = c[3] + t * S;
    S = c[2] + t * S;
    double dS = c[11];
    dS = c[10] + t * dS;
    dS = c[9] + t * dS;
    dS = c[8] + t * dS;
    dS = c[7] + t * dS;
    dS *= inv_rc;
    const double inv_4pi_r = 0.079577471545947673 / r;  // ESP_INV_4PI (ESPForceCompute.cc:36)
    const double inv_4pi_r2 = inv_4pi_r / r;
    L = S * inv_4pi_r;
    negdLdr = S * inv_4pi_r2 - dS * inv_4pi_r;
    }

template<typename Real> class EvaluatorPairPSWFT
    {
    public:
    struct param_type
        {
        Real kappa;                //!< splitting parameter [1/length] (diagnostics)
        Real alpha;                //!< Debye screening [1/length]; must be 0 (diagnostics)
        Real rcut;                 //!< real-space cutoff r_c of the split table
        Real V_cut;                //!< energy shift reference, usually 0 (L(r_c) = 0)
        const double* table_host;  //!< split table, host copy (CPU path)
        const double* table_device;  //!< split table, device copy (GPU path)
        unsigned int n_segs;       //!< segments of the table
        unsigned int pad_[3];

        ESP_EVAL_DEVICE void load_shared(char*& /*ptr*/, unsigned int& /*available_bytes*/) { }
        ESP_EVAL_HOSTDEVICE void allocate_shared(char*& /*ptr*/, unsigned int& /*available_bytes*/) const { }
        void set_memory_hints() const { }
        void set_memory_hint() const { }

#if !defined(__HIPCC__)
        param_type()
            : kappa(0), alpha(0), rcut(0), V_cut(0), table_host(nullptr), table_device(nullptr), n_segs(0), pad_ {0, 0, 0}
            {
            }
#if defined(ESP_EVALUATOR_PYBIND11)
        //! Keys "rcut", "kappa", "alpha", "V_cut" (EvaluatorPairPSWF.h:204-212) and the table's
        //! "table_ptr" (host address), "table_device_ptr" and "n_segs" (integers, from
        //! esp_get_pair_table).
        param_type(pybind11::dict v, bool /*managed*/ = false) : param_type()
            {
            rcut = v.contains("rcut") ? v["rcut"].cast<Real>() : Real(0);
            kappa = v.contains("kappa") ? v["kappa"].cast<Real>() : Real(0);
            alpha = v.contains("alpha") ? v["alpha"].cast<Real>() : Real(0);
            V_cut = v.contains("V_cut") ? v["V_cut"].cast<Real>() : Real(0);
            if (v.contains("table_ptr"))
                table_host = reinterpret_cast<const double*>(v["table_ptr"].cast<uintptr_t>());
            if (v.contains("table_device_ptr"))
                table_device = reinterpret_cast<const double*>(v["table_device_ptr"].cast<uintptr_t>());
            n_segs = v.contains("n_segs") ? v["n_segs"].cast<unsigned int>() : 0u;
            }

        pybind11::dict asDict() const
            {
            pybind11::dict v;
            v["kappa"] = kappa;
            v["alpha"] = alpha;
            v["rcut"] = rcut;
            v["V_cut"] = V_cut;
            v["table_ptr"] = reinterpret_cast<uintptr_t>(table_host);
            v["table_device_ptr"] = reinterpret_cast<uintptr_t>(table_device);
            v["n_segs"] = n_segs;
            return v;
            }
#endif
#endif
        } __attribute__((aligned(16)));

    ESP_EVAL_HOSTDEVICE EvaluatorPairPSWFT(Real _rsq, Real _rcutsq, const param_type& p)
        : rsq(_rsq), rcutsq(_rcutsq), rcut(p.rcut), V_cut(p.V_cut), qiqj(0), n_segs(p.n_segs)
        {
#if defined(__HIP_DEVICE_COMPILE__)
        table = p.table_device;
#else
        table = p.table_host;
#endif
        }

    ESP_EVAL_HOSTDEVICE static bool needsCharge() { return true; }
    ESP_EVAL_HOSTDEVICE void setCharge(Real qi, Real qj) { qiqj = qi * qj; }
    ESP_EVAL_HOSTDEVICE static bool needsDiameter() { return false; }
    ESP_EVAL_HOSTDEVICE void setDiameter(Real, Real) { }
    ESP_EVAL_HOSTDEVICE bool areInteractionParametersUseless() const { return rcut < Real(1.0e-12); }

    //! force_divr = F / r and the pair energy; false (outputs untouched) outside the cutoff
    //! or for a neutral pair.
    ESP_EVAL_HOSTDEVICE bool evalForceAndEnergy(Real& force_divr, Real& pair_eng, bool energy_shift) const
        {
        if (rsq >= rcutsq || qiqj == Real(0) || rcut < Real(1.0e-12))
            return false;
        const double r = sqrt(double(rsq));
        double L, negdLdr;
        pswf_lookup(table, n_segs, r, double(rcut), L, negdLdr);
        double e = double(qiqj) * L;
        if (energy_shift)
            e -= double(qiqj) * double(V_cut);
        pair_eng = Real(e);
        force_divr = Real(double(qiqj) * negdLdr / r);
        return true;
        }

    //! L(r) vanishes at r_c by construction: no tail corrections.
    ESP_EVAL_HOSTDEVICE Real evalPressureLRCIntegral() { return Real(0); }
    ESP_EVAL_HOSTDEVICE Real evalEnergyLRCIntegral() { return Real(0); }

#if !defined(__HIPCC__)
    static std::string getName() { return std::string("pswf"); }
    std::string getShapeSpec() const
        {
        throw std::runtime_error("EvaluatorPairPSWF: shape specification is not defined for a point-charge pair potential.");
        }
#endif

    protected:
    Real rsq, rcutsq, rcut, V_cut, qiqj;
    const double* table;
    unsigned int n_segs;
    };

//! The evaluator at HOOMD's double-precision Scalar (the reference build).
typedef EvaluatorPairPSWFT<double> EvaluatorPairPSWF;

} // namespace esp_hoomd

#endif /* ESP_EVALUATOR_PAIR_PSWF_H_ */
