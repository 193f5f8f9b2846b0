// Host-side mathematics of the ESP split; see pswf_math.h.
#include "pswf_math.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace esp
{
namespace math
{

void gauss_legendre(int n, double a, double b, std::vector<double>& x, std::vector<double>& w)
    {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int i = 0; i < (n + 1) / 2; ++i)
        {
        double z = std::cos(PI * (i + 0.75) / (n + 0.5));
        double dp = 1.0;
        for (int it = 0; it < 100; ++it)
            {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; ++k)
                {
                const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
                }
            if (n == 1)
                {
                p1 = z;
                p0 = 1.0;
                }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
            const double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16)
                break;
            }
        // recompute derivative at the converged node
        double p0 = 1.0, p1 = z;
        for (int k = 2; k <= n; ++k)
            {
            const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
            }
        dp = n * (z * p1 - p0) / (z * z - 1.0);
        const double wi = 2.0 / ((1.0 - z * z) * dp * dp);
        const double half = 0.5 * (b - a), mid = 0.5 * (b + a);
        x[i] = mid - half * z;
        x[n - 1 - i] = mid + half * z;
        w[i] = half * wi;
        w[n - 1 - i] = half * wi;
        }
    }

// ---------------------------------------------------------------------------
// Prolate0: psi_0^c = sum_m b_m sqrt(2m' + 1/2) P_{2m'}  (m' = m)
// The prolate operator -(1-x^2) d^2/dx^2 + 2x d/dx + c^2 x^2 is symmetric
// tridiagonal in the normalised even Legendre basis; psi_0 belongs to the
// smallest eigenvalue, found by Sturm-sequence bisection, the vector by
// inverse iteration.
// ---------------------------------------------------------------------------
Prolate0::Prolate0(double c) : m_c(c), m_chi(0.0)
    {
    if (!(c > 0.0) || c > 400.0)
        throw std::invalid_argument("Prolate0: bandwidth c must be in (0, 400]");
    const int n = std::max(24, int(std::ceil(c)) + 32);
    std::vector<double> a(n), e(n, 0.0);
    const double c2 = c * c;
    for (int m = 0; m < n; ++m)
        {
        const double k = 2.0 * m;
        a[m] = k * (k + 1.0) + c2 * (2.0 * k * (k + 1.0) - 1.0) / ((2.0 * k + 3.0) * (2.0 * k - 1.0));
        if (m + 1 < n)
            e[m] = c2 * (k + 2.0) * (k + 1.0) / ((2.0 * k + 3.0) * std::sqrt((2.0 * k + 1.0) * (2.0 * k + 5.0)));
        }
    // Gershgorin bounds
    double lo = a[0] - std::fabs(e[0]), hi = a[0] + std::fabs(e[0]);
    for (int m = 1; m < n; ++m)
        {
        const double r = std::fabs(e[m - 1]) + std::fabs(e[m]);
        lo = std::min(lo, a[m] - r);
        hi = std::max(hi, a[m] + r);
        }
    auto count_below = [&](double lam)
    {
        int cnt = 0;
        double d = a[0] - lam;
        if (d < 0)
            ++cnt;
        for (int m = 1; m < n; ++m)
            {
            if (d == 0.0)
                d = 1e-300;
            d = a[m] - lam - e[m - 1] * e[m - 1] / d;
            if (d < 0)
                ++cnt;
            }
        return cnt;
    };
    for (int it = 0; it < 200; ++it)
        {
        const double mid = 0.5 * (lo + hi);
        if (count_below(mid) >= 1)
            hi = mid;
        else
            lo = mid;
        if (hi - lo <= 4e-16 * std::max(1.0, std::fabs(hi)))
            break;
        }
    m_chi = 0.5 * (lo + hi);
    // inverse iteration with a shift just below chi
    const double sigma = m_chi - 1e-10 * std::max(1.0, std::fabs(m_chi));
    std::vector<double> v(n, 1.0), y(n), cp(n), dp(n);
    for (int it = 0; it < 4; ++it)
        {
        // Thomas solve (A - sigma) y = v
        double den = a[0] - sigma;
        cp[0] = e[0] / den;
        dp[0] = v[0] / den;
        for (int m = 1; m < n; ++m)
            {
            den = a[m] - sigma - e[m - 1] * cp[m - 1];
            cp[m] = (m + 1 < n) ? e[m] / den : 0.0;
            dp[m] = (v[m] - e[m - 1] * dp[m - 1]) / den;
            }
        y[n - 1] = dp[n - 1];
        for (int m = n - 2; m >= 0; --m)
            y[m] = dp[m] - cp[m] * y[m + 1];
        double nrm = 0.0;
        for (double t : y)
            nrm += t * t;
        nrm = std::sqrt(nrm);
        for (int m = 0; m < n; ++m)
            v[m] = y[m] / nrm;
        }
    if (v[0] < 0)
        for (double& t : v)
            t = -t;
    m_b = v;
    }

double Prolate0::eval(double x) const
    {
    // sum_m b_m sqrt(2m+1/2) P_{2m}(x) with the three-term recurrence
    double p0 = 1.0, p1 = x;
    double s = m_b[0] * std::sqrt(0.5);
    const int n = int(m_b.size());
    int deg = 1;
    for (int m = 1; m < n; ++m)
        {
        const int k = 2 * m;
        while (deg < k)
            {
            const double p2 = ((2.0 * deg + 1.0) * x * p1 - deg * p0) / (deg + 1.0);
            p0 = p1;
            p1 = p2;
            ++deg;
            }
        s += m_b[m] * std::sqrt(k + 0.5) * p1;
        }
    return s;
    }

double Prolate0::integral(double x) const
    {
    // int_0^x P_0 = x; int_0^x P_k = (P_{k+1}(x) - P_{k-1}(x)) / (2k+1) for even k >= 2
    const int n = int(m_b.size());
    const int kmax = 2 * (n - 1) + 1;
    std::vector<double> P(kmax + 1);
    P[0] = 1.0;
    if (kmax >= 1)
        P[1] = x;
    for (int d = 1; d < kmax; ++d)
        P[d + 1] = ((2.0 * d + 1.0) * x * P[d] - d * P[d - 1]) / (d + 1.0);
    double s = m_b[0] * std::sqrt(0.5) * x;
    for (int m = 1; m < n; ++m)
        {
        const int k = 2 * m;
        s += m_b[m] * std::sqrt(k + 0.5) * (P[k + 1] - P[k - 1]) / (2.0 * k + 1.0);
        }
    return s;
    }

// ---------------------------------------------------------------------------
// SplitFunction
// ---------------------------------------------------------------------------
SplitFunction::SplitFunction(SplitKind kind, double kappa, double rc)
    : m_kind(kind), m_kappa(kappa), m_rc(rc), m_c(kappa * rc), m_Z(1.0), m_psi(nullptr)
    {
    if (!(rc > 0.0))
        throw std::invalid_argument("SplitFunction: rcut must be > 0");
    if (!(kappa > 0.0))
        throw std::invalid_argument("SplitFunction: kappa must be > 0");
    if (kind == SplitKind::PSWF)
        {
        m_psi = new Prolate0(m_c);
        m_Z = m_psi->integral(1.0);
        // quadrature nodes on [0, 1] for moments and F at moderate k
        gauss_legendre(quadrature_nodes(m_c, 4.0 * m_c + 64.0), 0.0, 1.0, m_qx, m_qw);
        }
    }

SplitFunction::~SplitFunction()
    {
    delete m_psi;
    }

double SplitFunction::S(double s) const
    {
    if (m_kind == SplitKind::GAUSSIAN)
        return std::erfc(m_kappa * m_rc * s);
    if (s >= 1.0)
        return 0.0;
    if (s <= 0.0)
        return 1.0;
    return 1.0 - m_psi->integral(s) / m_Z;
    }

double SplitFunction::dSds(double s) const
    {
    if (m_kind == SplitKind::GAUSSIAN)
        {
        const double a = m_kappa * m_rc;
        return -2.0 * a / std::sqrt(PI) * std::exp(-a * a * s * s);
        }
    if (s > 1.0)
        return 0.0;
    return -m_psi->eval(std::max(s, 0.0)) / m_Z;
    }

int SplitFunction::quadrature_nodes(double c, double omega_max)
    {
    const int n = int(std::ceil(0.5 * omega_max + 0.5 * c)) + 40;
    return std::min(std::max(n, 48), 1024);
    }

void SplitFunction::quadrature(int n, std::vector<double>& x, std::vector<double>& wpsi) const
    {
    if (m_kind != SplitKind::PSWF)
        throw std::logic_error("SplitFunction::quadrature: PSWF split only");
    std::vector<double> w;
    gauss_legendre(n, 0.0, 1.0, x, w);
    wpsi.resize(n);
    for (int i = 0; i < n; ++i)
        wpsi[i] = w[i] * m_psi->eval(x[i]) / m_Z;
    }

double SplitFunction::F(double k) const
    {
    if (m_kind == SplitKind::GAUSSIAN)
        return std::exp(-k * k / (4.0 * m_kappa * m_kappa));
    std::vector<double> x, w;
    const int n = quadrature_nodes(m_c, std::fabs(k) * m_rc);
    gauss_legendre(n, 0.0, 1.0, x, w);
    double s = 0.0;
    for (int i = 0; i < n; ++i)
        s += w[i] * m_psi->eval(x[i]) * std::cos(k * m_rc * x[i]);
    return s / m_Z;
    }

double SplitFunction::dF(double k) const
    {
    if (m_kind == SplitKind::GAUSSIAN)
        return -k / (2.0 * m_kappa * m_kappa) * F(k);
    std::vector<double> x, w;
    const int n = quadrature_nodes(m_c, std::fabs(k) * m_rc);
    gauss_legendre(n, 0.0, 1.0, x, w);
    double s = 0.0;
    for (int i = 0; i < n; ++i)
        s -= w[i] * m_psi->eval(x[i]) * m_rc * x[i] * std::sin(k * m_rc * x[i]);
    return s / m_Z;
    }

double SplitFunction::self_coeff() const
    {
    if (m_kind == SplitKind::GAUSSIAN)
        return 2.0 * m_kappa / std::sqrt(PI);
    return m_psi->eval(0.0) / (m_Z * m_rc);
    }

double SplitFunction::background_coeff() const
    {
    if (m_kind == SplitKind::GAUSSIAN)
        return 1.0 / (4.0 * m_kappa * m_kappa);
    double s = 0.0;
    for (size_t i = 0; i < m_qx.size(); ++i)
        s += m_qw[i] * m_qx[i] * m_qx[i] * m_psi->eval(m_qx[i]);
    return m_rc * m_rc * s / (2.0 * m_Z);
    }

// ---------------------------------------------------------------------------
// Polynomial fitting helpers
// ---------------------------------------------------------------------------
namespace
{
//! Chebyshev interpolation of f on z in [-1,1] at the D+1 Chebyshev-Lobatto
//! points, returned as monomial coefficients in z.
template<class Fn> std::vector<double> cheb_fit_monomial(unsigned int D, Fn f)
    {
    const unsigned int n = D + 1;
    std::vector<double> fz(n), cheb(n, 0.0);
    for (unsigned int j = 0; j < n; ++j)
        {
        const double z = (D == 0) ? 0.0 : std::cos(PI * double(j) / double(D));
        fz[j] = f(z);
        }
    if (D == 0)
        return {fz[0]};
    // Chebyshev coefficients from Lobatto samples (discrete cosine transform, type I)
    for (unsigned int k = 0; k < n; ++k)
        {
        double s = 0.0;
        for (unsigned int j = 0; j < n; ++j)
            {
            const double wgt = (j == 0 || j == D) ? 0.5 : 1.0;
            s += wgt * fz[j] * std::cos(PI * double(k) * double(j) / double(D));
            }
        cheb[k] = 2.0 * s / double(D);
        }
    cheb[0] *= 0.5;
    cheb[D] *= 0.5;
    // convert sum_k cheb[k] T_k(z) to monomials
    std::vector<double> mono(n, 0.0), tkm1(n, 0.0), tk(n, 0.0), tkp1(n, 0.0);
    tkm1[0] = 1.0;  // T_0
    tk[1] = 1.0;    // T_1
    for (unsigned int j = 0; j < n; ++j)
        mono[j] += cheb[0] * tkm1[j];
    for (unsigned int j = 0; j < n; ++j)
        mono[j] += cheb[1] * tk[j];
    for (unsigned int k = 2; k < n; ++k)
        {
        std::fill(tkp1.begin(), tkp1.end(), 0.0);
        for (unsigned int j = 0; j + 1 < n; ++j)
            tkp1[j + 1] += 2.0 * tk[j];
        for (unsigned int j = 0; j < n; ++j)
            tkp1[j] -= tkm1[j];
        for (unsigned int j = 0; j < n; ++j)
            mono[j] += cheb[k] * tkp1[j];
        tkm1.swap(tk);
        tk.swap(tkp1);
        }
    return mono;
    }

//! Degree-4 interpolation in t in [0,1] at Chebyshev-Lobatto points, returned in t-monomials.
template<class Fn> void fit_segment_t(Fn f, double* out)
    {
    // fit in z = 2t - 1, then expand p(2t - 1) into powers of t
    std::vector<double> mz = cheb_fit_monomial(4, [&](double z) { return f(0.5 * (z + 1.0)); });
    double mt[5] = {0, 0, 0, 0, 0};
    // (2t - 1)^j = sum_i binom(j,i) 2^i t^i (-1)^(j-i)
    const double binom[5][5] = {{1, 0, 0, 0, 0}, {1, 1, 0, 0, 0}, {1, 2, 1, 0, 0}, {1, 3, 3, 1, 0}, {1, 4, 6, 4, 1}};
    for (int j = 0; j < 5; ++j)
        for (int i = 0; i <= j; ++i)
            mt[i] += mz[j] * binom[j][i] * std::pow(2.0, i) * (((j - i) % 2) ? -1.0 : 1.0);
    for (int i = 0; i < 5; ++i)
        out[i] = mt[i];
    }
} // namespace

void build_split_table(const SplitFunction& split,
                       unsigned int n_segs,
                       std::vector<double>& screen,
                       std::vector<double>& dscreen,
                       std::vector<double>& r_lo,
                       std::vector<double>& dr_inv)
    {
    if (n_segs < 1)
        throw std::invalid_argument("build_split_table: n_segs must be >= 1");
    screen.assign(size_t(n_segs) * 5, 0.0);
    dscreen.assign(size_t(n_segs) * 5, 0.0);
    r_lo.resize(n_segs);
    dr_inv.resize(n_segs);
    const double ds = 1.0 / double(n_segs);
    for (unsigned int i = 0; i < n_segs; ++i)
        {
        const double s0 = double(i) * ds;
        r_lo[i] = s0;
        dr_inv[i] = double(n_segs);
        fit_segment_t([&](double t) { return split.S(s0 + t * ds); }, &screen[size_t(i) * 5]);
        fit_segment_t([&](double t) { return split.dSds(s0 + t * ds); }, &dscreen[size_t(i) * 5]);
        }
    }

void build_compact_split(const SplitFunction& split, unsigned int nseg, unsigned int deg, std::vector<double>& coef)
    {
    if (nseg < 1 || deg < 1 || deg > 15)
        throw std::invalid_argument("build_compact_split: invalid segment count or degree");
    coef.assign(size_t(nseg) * (deg + 1), 0.0);
    for (unsigned int i = 0; i < nseg; ++i)
        {
        std::vector<double> m = cheb_fit_monomial(deg,
                                                  [&](double z)
                                                  {
                                                      const double s = (double(i) + 0.5 * (z + 1.0)) / double(nseg);
                                                      return split.S(s);
                                                  });
        for (unsigned int j = 0; j <= deg; ++j)
            coef[size_t(i) * (deg + 1) + j] = m[j];
        }
    }

// ---------------------------------------------------------------------------
// Spreading window
// ---------------------------------------------------------------------------
unsigned int default_window_degree(unsigned int P)
    {
    return std::min(P + 3u, 15u);
    }

double default_window_c(unsigned int P, double sigma)
    {
    // near-optimal prolate bandwidth for a P-point window at oversampling sigma
    const double s = std::min(std::max(sigma, 1.05), 3.0);
    return PI * double(P) * (1.0 - 0.5 / s);
    }

double WindowTable::eval(unsigned int k, double t) const
    {
    const double z = 2.0 * t - 1.0;
    const double* c = &coef[size_t(k) * (D + 1)];
    double v = c[D];
    for (int j = int(D) - 1; j >= 0; --j)
        v = v * z + c[j];
    return v;
    }

double WindowTable::fourier(double theta) const
    {
    // int W(d) exp(-i theta d) dd over the P pieces; 24-point Gauss per piece is exact to
    // rounding for |theta| <= pi and D <= 15
    static thread_local std::vector<double> gx, gw;
    if (gx.size() != 24)
        gauss_legendre(24, 0.0, 1.0, gx, gw);
    double re = 0.0, im = 0.0;
    for (unsigned int k = 0; k < P; ++k)
        for (size_t q = 0; q < gx.size(); ++q)
            {
            const double t = gx[q];
            const double d = double(k) - 0.5 * double(P) + t;
            const double w = gw[q] * eval(k, t);
            re += w * std::cos(theta * d);
            im -= w * std::sin(theta * d);
            }
    return std::sqrt(re * re + im * im);
    }

WindowTable build_window(unsigned int P, unsigned int D, double cw)
    {
    if (P < 2 || P > 16)
        throw std::invalid_argument("build_window: order P must be in [2, 16]");
    if (D < 2 || D > 15)
        throw std::invalid_argument("build_window: window degree must be in [2, 15]");
    Prolate0 psi(cw);
    const double norm = double(P) * psi.integral(1.0);  // int_{-P/2}^{P/2} psi(2d/P) dd
    WindowTable w;
    w.P = P;
    w.D = D;
    w.cw = cw;
    w.coef.assign(size_t(P) * (D + 1), 0.0);
    for (unsigned int k = 0; k < P; ++k)
        {
        auto f = [&](double z)
        {
            const double t = 0.5 * (z + 1.0);
            const double d = double(k) - 0.5 * double(P) + t;
            const double u = std::min(1.0, std::max(-1.0, 2.0 * d / double(P)));
            return psi.eval(u) / norm;
        };
        std::vector<double> m = cheb_fit_monomial(D, f);
        for (unsigned int j = 0; j <= D; ++j)
            w.coef[size_t(k) * (D + 1) + j] = m[j];
        }
    // symmetrise exactly: tap P-1-k at 1-t mirrors tap k at t (z -> -z)
    for (unsigned int k = 0; k < P / 2; ++k)
        {
        const unsigned int km = P - 1 - k;
        for (unsigned int j = 0; j <= D; ++j)
            {
            const double a = w.coef[size_t(k) * (D + 1) + j];
            const double b = w.coef[size_t(km) * (D + 1) + j] * ((j % 2) ? -1.0 : 1.0);
            const double avg = 0.5 * (a + b);
            w.coef[size_t(k) * (D + 1) + j] = avg;
            w.coef[size_t(km) * (D + 1) + j] = avg * ((j % 2) ? -1.0 : 1.0);
            }
        }
    // renormalise the fitted pieces exactly: int_0^1 z^j dt = 1/(j+1) for even j, 0 for odd j
    double total = 0.0;
    for (unsigned int k = 0; k < P; ++k)
        for (unsigned int j = 0; j <= D; j += 2)
            total += w.coef[size_t(k) * (D + 1) + j] / double(j + 1);
    for (double& c : w.coef)
        c /= total;
    return w;
    }

    } // namespace math
    } // namespace esp
