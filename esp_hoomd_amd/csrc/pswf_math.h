// Host-side mathematics of the ESP (Ewald summation with prolates) split.
//
// The Coulomb kernel 1/(4 pi r) is split as L(r) + M(r):
//
//   L(r)  = S(r / r_c) / (4 pi r)                 short range, exactly 0 for r >= r_c
//   S(s)  = 1 - (1/Z) * int_0^s psi_0^c(x) dx      S(0) = 1, S(1) = 0
//   dS/ds = -psi_0^c(s) / Z,   Z = int_0^1 psi_0^c(x) dx
//   M^(k) = F(k) / k^2,  F(k) = (1/Z) int_0^1 psi_0^c(x) cos(k r_c x) dx,  F(0) = 1
//
// psi_0^c is the zeroth prolate spheroidal wave function with bandwidth c,
// computed from its expansion in normalised Legendre polynomials (the
// eigenvector of the prolate differential operator, which is tridiagonal in
// that basis).  The same functions give the charge-spreading window
// W(d) ~ psi_0^{c_w}(2 d / P) on the P-point stencil.
//
// Table formats follow the reference's PSWF_Coeffs.h: per segment five
// Horner coefficients c0..c4 in t = (s - r_lo[seg]) * dr_inv[seg]
// (PSWF_Coeffs.h:24-9599), here in the reduced variable s = r / r_c.
#pragma once

#include <cstddef>
#include <vector>

namespace esp
{
namespace math
{

constexpr double INV_4PI = 0.079577471545947673;  // ESPForceCompute.cc:36
constexpr double PI = 3.14159265358979323846;

//! Gauss-Legendre nodes and weights on [a, b].
void gauss_legendre(int n, double a, double b, std::vector<double>& x, std::vector<double>& w);

//! psi_0^c on [-1, 1] (and its analytic continuation), sign fixed by psi(0) > 0.
class Prolate0
    {
    public:
    explicit Prolate0(double c);
    double c() const { return m_c; }
    double eval(double x) const;
    //! int_0^x psi_0^c(t) dt
    double integral(double x) const;
    //! smallest eigenvalue chi_0 of the prolate operator (diagnostic)
    double chi() const { return m_chi; }

    private:
    double m_c;
    double m_chi;
    std::vector<double> m_b;  // coefficients of sqrt(2k+1/2) P_{2k}
    };

//! Split kinds: PSWF (the ESP split) or Gaussian (erfc; textbook Ewald/SPME, used to validate).
enum class SplitKind
    {
    PSWF = 0,
    GAUSSIAN = 1
    };

//! The radial split function and its Fourier-space complement.
class SplitFunction
    {
    public:
    //! PSWF: bandwidth c = kappa * rc.  Gaussian: S = erfc(kappa r).
    SplitFunction(SplitKind kind, double kappa, double rc);
    SplitKind kind() const { return m_kind; }
    double rc() const { return m_rc; }
    double kappa() const { return m_kappa; }
    double S(double s) const;      //!< reduced s = r / rc
    double dSds(double s) const;   //!< dS/ds
    //! F(k) and dF/dk of the long-range part, F(0) = 1
    double F(double k) const;
    double dF(double k) const;
    //! lim_{r->0} (1 - S(r/rc)) / r = -dS/dr(0); the self energy is -INV_4PI/2 * this * sum q^2
    double self_coeff() const;
    //! lim_{k->0} (1 - F(k)) / k^2; the neutralising-background energy is -Q^2/(2V) * this
    double background_coeff() const;
    //! quadrature data for the device influence kernel: F(k) = sum_i wpsi[i] cos(k rc x[i])
    void quadrature(int n, std::vector<double>& x, std::vector<double>& wpsi) const;
    //! number of quadrature nodes needed for |k| rc <= omega_max
    static int quadrature_nodes(double c, double omega_max);

    private:
    SplitKind m_kind;
    double m_kappa;
    double m_rc;
    double m_c;
    double m_Z;
    Prolate0* m_psi;  // owned; PSWF only
    std::vector<double> m_qx, m_qw;  // cached quadrature for F / dF / moments

    public:
    ~SplitFunction();
    SplitFunction(const SplitFunction&) = delete;
    SplitFunction& operator=(const SplitFunction&) = delete;
    };

//! Build the reference-format split table: n_segs x 5 Horner coefficients for S and for dS/ds,
//! segment bounds r_lo[i] = i / n_segs and dr_inv[i] = n_segs (reduced units).
void build_split_table(const SplitFunction& split,
                       unsigned int n_segs,
                       std::vector<double>& screen,
                       std::vector<double>& dscreen,
                       std::vector<double>& r_lo,
                       std::vector<double>& dr_inv);

//! Compact device table of S(s): nseg segments, each a degree-deg polynomial in z = 2t - 1
//! (t the offset inside the segment), coef[seg * (deg + 1) + j] for z^j.
void build_compact_split(const SplitFunction& split, unsigned int nseg, unsigned int deg, std::vector<double>& coef);

//! P-tap piecewise-polynomial spreading window.  Tap k (0 <= k < P) covers grid offsets
//! d = k - P/2 + t, t in [0, 1); its weight is a polynomial of degree D in z = 2t - 1,
//! stored as coef[k * (D + 1) + j] for z^j.  Normalised so that int W(d) dd = 1.
struct WindowTable
    {
    unsigned int P = 0;
    unsigned int D = 0;
    double cw = 0.0;
    std::vector<double> coef;
    //! continuous Fourier transform of the piecewise polynomial at grid-unit frequency theta
    double fourier(double theta) const;
    double eval(unsigned int k, double t) const;
    };

WindowTable build_window(unsigned int P, unsigned int D, double cw);

//! Default window degree and bandwidth for a P-point stencil with oversampling sigma.
unsigned int default_window_degree(unsigned int P);
double default_window_c(unsigned int P, double sigma);

    } // namespace math
    } // namespace esp
