// Host orchestration of one ESP solve; see esp_solver.h.
#include "esp_solver.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <tuple>

#include "esp_kernels.h"

namespace esp
{

// ---------------------------------------------------------------------------
// Process-wide cache of the host tables.  In the reduced variable s = r / r_c the split
// tables, the compact pair table and the influence quadrature depend on the bandwidth
// c = kappa r_c only (the self and background constants scale as 1/r_c and r_c^2), and the
// window on (P, D, c_w) only; many solvers of one batch or server share them, and building
// them (prolate eigenproblem, quadratures, fits) costs tens of milliseconds.
// ---------------------------------------------------------------------------
namespace
{
struct ReducedSplit
    {
    std::vector<double> screen, dscreen, r_lo, dr_inv;
    double self_rc = 0.0;     // self_coeff * r_c
    double bg_rc2 = 0.0;      // background_coeff / r_c^2
    };

std::mutex g_cache_mutex;
std::map<std::tuple<int, double, unsigned int>, std::shared_ptr<const ReducedSplit>> g_split_cache;
std::map<std::tuple<int, double, unsigned int, unsigned int>, std::shared_ptr<const std::vector<double>>> g_compact_cache;
std::map<std::tuple<unsigned int, unsigned int, double>, std::shared_ptr<const math::WindowTable>> g_window_cache;
std::map<std::tuple<int, double, int>, std::shared_ptr<const std::pair<std::vector<double>, std::vector<double>>>> g_quad_cache;

std::shared_ptr<const ReducedSplit> reduced_split(math::SplitKind kind, double c, unsigned int n_table)
    {
    const auto key = std::make_tuple(int(kind), c, n_table);
    std::lock_guard<std::mutex> lock(g_cache_mutex);
    auto it = g_split_cache.find(key);
    if (it != g_split_cache.end())
        return it->second;
    auto t = std::make_shared<ReducedSplit>();
    math::SplitFunction split(kind, c, 1.0);  // r_c = 1: reduced units
    math::build_split_table(split, n_table, t->screen, t->dscreen, t->r_lo, t->dr_inv);
    t->self_rc = split.self_coeff();
    t->bg_rc2 = split.background_coeff();
    g_split_cache.emplace(key, t);
    return t;
    }

std::shared_ptr<const std::vector<double>> compact_split(math::SplitKind kind, double c, unsigned int nseg, unsigned int deg)
    {
    const auto key = std::make_tuple(int(kind), c, nseg, deg);
    std::lock_guard<std::mutex> lock(g_cache_mutex);
    auto it = g_compact_cache.find(key);
    if (it != g_compact_cache.end())
        return it->second;
    math::SplitFunction split(kind, c, 1.0);
    auto t = std::make_shared<std::vector<double>>();
    math::build_compact_split(split, nseg, deg, *t);
    g_compact_cache.emplace(key, t);
    return t;
    }

std::shared_ptr<const math::WindowTable> cached_window(unsigned int P, unsigned int D, double cw)
    {
    const auto key = std::make_tuple(P, D, cw);
    std::lock_guard<std::mutex> lock(g_cache_mutex);
    auto it = g_window_cache.find(key);
    if (it != g_window_cache.end())
        return it->second;
    auto w = std::make_shared<math::WindowTable>(math::build_window(P, D, cw));
    g_window_cache.emplace(key, w);
    return w;
    }

//! Quadrature (x in [0, 1], weights) with F(k) = sum w cos(k r_c x): depends on c and the node count.
std::shared_ptr<const std::pair<std::vector<double>, std::vector<double>>> cached_quadrature(math::SplitKind kind, double c, int nq)
    {
    const auto key = std::make_tuple(int(kind), c, nq);
    std::lock_guard<std::mutex> lock(g_cache_mutex);
    auto it = g_quad_cache.find(key);
    if (it != g_quad_cache.end())
        return it->second;
    math::SplitFunction split(kind, c, 1.0);
    auto q = std::make_shared<std::pair<std::vector<double>, std::vector<double>>>();
    split.quadrature(nq, q->first, q->second);
    g_quad_cache.emplace(key, q);
    return q;
    }
} // namespace

void check_hip(hipError_t e, const char* what)
    {
    if (e != hipSuccess)
        {
        std::ostringstream oss;
        oss << what << ": " << hipGetErrorString(e);
        throw EspError(e == hipErrorOutOfMemory ? STATUS_OOM : STATUS_HIP, oss.str());
        }
    }

DeviceBuffer::~DeviceBuffer()
    {
    release();
    }

void DeviceBuffer::release()
    {
    if (m_ptr)
        (void)hipFree(m_ptr);
    m_ptr = nullptr;
    m_bytes = 0;
    }

void DeviceBuffer::resize(size_t bytes, bool zero)
    {
    if (bytes > m_bytes)
        {
        release();
        if (bytes > 0)
            {
            check_hip(hipMalloc(&m_ptr, bytes), "hipMalloc");
            m_bytes = bytes;
            }
        if (zero && m_ptr)
            check_hip(hipMemset(m_ptr, 0, m_bytes), "hipMemset");
        }
    else if (zero && m_ptr)
        check_hip(hipMemset(m_ptr, 0, m_bytes), "hipMemset");
    }

void DeviceBuffer::resize_zeroed_async(size_t bytes, hipStream_t s)
    {
    resize(bytes, false);
    if (m_ptr)
        check_hip(hipMemsetAsync(m_ptr, 0, m_bytes, s), "hipMemsetAsync");
    }

FFTPlan make_fft_plan(int n)
    {
    FFTPlan p {};
    p.n = n;
    p.nstages = 0;
    int m = n;
    if (n < 1)
        throw EspError(STATUS_ARG, "mesh dimension must be >= 1");
    while (m % 4 == 0 && p.nstages < MAX_FFT_STAGES)
        {
        p.radix[p.nstages++] = 4;
        m /= 4;
        }
    const int rads[3] = {2, 3, 5};
    for (int r : rads)
        while (m % r == 0 && p.nstages < MAX_FFT_STAGES)
            {
            p.radix[p.nstages++] = r;
            m /= r;
            }
    if (m != 1)
        {
        std::ostringstream oss;
        oss << "mesh dimension " << n << " has a prime factor other than 2, 3, 5";
        throw EspError(STATUS_UNSUPPORTED, oss.str());
        }
    return p;
    }

EspSolver::EspSolver() = default;

EspSolver::~EspSolver()
    {
    release_graph();
    if (m_stage_ev)
        {
        (void)hipEventSynchronize(m_stage_ev);
        (void)hipEventDestroy(m_stage_ev);
        }
    if (m_stage)
        (void)hipHostFree(m_stage);
    for (unsigned char* p : m_stage_retired)
        (void)hipHostFree(p);
    if (m_gstream)
        (void)hipStreamDestroy(m_gstream);
    for (auto& v : m_ev_pool)
        for (hipEvent_t e : v)
            (void)hipEventDestroy(e);
    for (auto& v : m_ev_pending)
        for (hipEvent_t e : v)
            (void)hipEventDestroy(e);
    if (m_side)
        {
        (void)hipStreamSynchronize(m_side);
        (void)hipStreamDestroy(m_side);
        (void)hipEventDestroy(m_ev_fork);
        (void)hipEventDestroy(m_ev_join);
        (void)hipEventDestroy(m_ev_zdone);
        }
    }

const char* EspSolver::stage_name(int i)
    {
    static const char* names[N_STAGES] = {"pair",   "bin",          "scan",       "scatter", "spread",  "fft_x_r2c",
                                          "fft_y", "fft_z_convolve", "ifft_y", "ifft_x_c2r", "gather", "finalize"};
    return (i >= 0 && i < N_STAGES) ? names[i] : "";
    }

void EspSolver::set_profiling(int level)
    {
    ++m_gen;
    m_profile = (level >= 0 && level <= 2) ? level : 1;
    }

void EspSolver::reset_stage_times()
    {
    harvest();
    for (int i = 0; i < N_STAGES; ++i)
        {
        m_stage_ms[i] = 0.0;
        m_stage_calls[i] = 0;
        }
    }

//! Record mesh-pipeline boundary b (0..N_STAGES) of the current call on stream s; stage i in
//! 1..N_STAGES-1 spans boundaries i and i+1.  Boundary 0 opens a new event set, N_STAGES closes
//! it.  At level 2 no boundary is recorded (only the pair stage's own events, on its stream).
void EspSolver::mark(int b, hipStream_t s)
    {
    if (!m_profile)
        return;
    if (m_profile == 2 && b > 0 && b < N_STAGES)
        return;
    if (b == 0)
        {
        if (m_ev_pool.empty())
            {
            std::vector<hipEvent_t> v(N_STAGES + 3);
            for (hipEvent_t& e : v)
                check_hip(hipEventCreate(&e), "hipEventCreate");
            m_ev_pool.push_back(std::move(v));
            }
        m_ev_pending.push_back(std::move(m_ev_pool.back()));
        m_ev_pool.pop_back();
        m_ev_marked.push_back(false);
        m_ev_full.push_back(m_profile == 1);
        m_ev_cur = &m_ev_pending.back();
        }
    if (!m_ev_cur)
        return;
    if (m_profile == 1)
        check_hip(hipEventRecord((*m_ev_cur)[b], s), "hipEventRecord");
    if (b == N_STAGES)
        m_ev_cur = nullptr;
    }

//! Record the start (end = false) or the end of the pair stage on the stream it runs on.
void EspSolver::pair_mark(bool end, hipStream_t s)
    {
    if (!m_profile || !m_ev_cur)
        return;
    check_hip(hipEventRecord((*m_ev_cur)[N_STAGES + (end ? 2 : 1)], s), "hipEventRecord");
    if (end)
        m_ev_marked.back() = true;
    }

void EspSolver::harvest()
    {
    for (size_t c = 0; c < m_ev_pending.size(); ++c)
        {
        std::vector<hipEvent_t>& v = m_ev_pending[c];
        if (m_ev_full[c])
            check_hip(hipEventSynchronize(v[N_STAGES]), "hipEventSynchronize");
        float ms = 0.0f;
        if (m_ev_marked[c])
            {
            check_hip(hipEventSynchronize(v[N_STAGES + 2]), "hipEventSynchronize");
            if (hipEventElapsedTime(&ms, v[N_STAGES + 1], v[N_STAGES + 2]) == hipSuccess)
                {
                m_stage_ms[0] += ms;
                m_stage_calls[0] += 1;
                }
            }
        if (m_ev_full[c])
            for (int i = 1; i < N_STAGES; ++i)
                if (hipEventElapsedTime(&ms, v[i], v[i + 1]) == hipSuccess)
                    {
                    m_stage_ms[i] += ms;
                    m_stage_calls[i] += 1;
                    }
        m_ev_pool.push_back(std::move(v));
        }
    m_ev_pending.clear();
    m_ev_marked.clear();
    m_ev_full.clear();
    }

void EspSolver::stage_times(double* ms, unsigned int* calls)
    {
    harvest();
    for (int i = 0; i < N_STAGES; ++i)
        {
        ms[i] = m_stage_ms[i];
        calls[i] = m_stage_calls[i];
        }
    }

void EspSolver::set_params(const SolverParams& p)
    {
    ++m_gen;
    if (p.nx < 2 || p.ny < 2 || p.nz < 2)
        throw EspError(STATUS_ARG, "mesh resolution must be >= 2 along each axis");
    if (p.nx > 1024 || p.ny > 1024 || p.nz > 1024)
        throw EspError(STATUS_ARG, "mesh resolution must be <= 1024 along each axis");
    if (p.order < 2 || p.order > unsigned(MAX_ORDER))
        throw EspError(STATUS_ARG, "order must be in [2, 16]");
    if (!(p.rcut > 0.0))
        throw EspError(STATUS_ARG, "rcut must be > 0");
    if (!(p.kappa > 0.0))
        throw EspError(STATUS_ARG, "kappa must be > 0");
    if (!(p.alpha >= 0.0) || !std::isfinite(p.alpha))
        throw EspError(STATUS_ARG, "alpha must be finite and >= 0");
    if (p.n_table < 16 || p.n_table > (1u << 20))
        throw EspError(STATUS_ARG, "n_table must be in [16, 2^20]");
    if (p.precision != 0 && p.precision != 1)
        throw EspError(STATUS_ARG, "precision must be 0 (FP64) or 1 (mixed)");
    if (p.split != 0 && p.split != 1)
        throw EspError(STATUS_ARG, "split must be 0 (PSWF) or 1 (Gaussian)");
    if (p.split == 0 && p.kappa * p.rcut > 200.0)
        throw EspError(STATUS_ARG, "PSWF bandwidth kappa * rcut must be <= 200");
    if (p.window_degree != 0 && (p.window_degree < 2 || p.window_degree > unsigned(MAX_WINDOW_DEG)))
        throw EspError(STATUS_ARG, "window_degree must be 0 or in [2, 15]");
    if (p.pair_lanes != 0 && p.pair_lanes != 1 && p.pair_lanes != 2 && p.pair_lanes != 4 && p.pair_lanes != 8 && p.pair_lanes != 16
        && p.pair_lanes != 32)
        throw EspError(STATUS_ARG, "pair_lanes must be 0 or a power of two <= 32");
    if (!std::isfinite(p.k_cut))
        throw EspError(STATUS_ARG, "k_cut must be finite");
    if (p.tile != 0 && (p.tile < 2 || p.tile > 16))
        throw EspError(STATUS_ARG, "tile must be 0 (automatic) or an edge from 2 to 16");
    // validate FFT factorisation eagerly
    make_fft_plan(int(p.nx));
    make_fft_plan(int(p.ny));
    make_fft_plan(int(p.nz));
    m_p = p;
    m_need_init = true;
    m_tables_built = false;
    m_external_table = false;
    }

void EspSolver::set_split_table(unsigned int n_segs, const double* screen, const double* dscreen, const double* r_lo, const double* dr_inv)
    {
    ++m_gen;
    if (n_segs < 1 || !screen || !dscreen || !r_lo || !dr_inv)
        throw EspError(STATUS_ARG, "set_split_table: invalid table");
    m_screen.assign(screen, screen + size_t(n_segs) * 5);
    m_dscreen.assign(dscreen, dscreen + size_t(n_segs) * 5);
    m_r_lo.assign(r_lo, r_lo + n_segs);
    m_dr_inv.assign(dr_inv, dr_inv + n_segs);
    m_external_table = true;
    m_need_init = true;
    }

void EspSolver::build_host_tables()
    {
    const math::SplitKind kind = m_p.split == 1 ? math::SplitKind::GAUSSIAN : math::SplitKind::PSWF;
    const auto red = reduced_split(kind, m_p.kappa * m_p.rcut, m_p.n_table);
    if (!m_external_table)
        {
        m_screen = red->screen;
        m_dscreen = red->dscreen;
        m_r_lo = red->r_lo;
        m_dr_inv = red->dr_inv;
        }
    m_self_coeff = red->self_rc / m_p.rcut;
    m_bg_coeff = red->bg_rc2 * m_p.rcut * m_p.rcut;

    // window: oversampling sigma = (pi / h) / K, K the effective band limit of F
    const unsigned int P = m_p.order;
    m_D = m_p.window_degree ? m_p.window_degree : math::default_window_degree(P);
    double cw = m_p.window_c;
    if (cw <= 0.0)
        {
        // mesh spacing is only known with the box; use the coarsest axis of the last box if any
        // (the length of lattice vector a over its mesh points: a tilted axis is longer than L)
        const double a1 = m_box[0], a2 = m_box[1] * std::sqrt(1.0 + m_box[3] * m_box[3]),
                     a3 = m_box[2] * std::sqrt(1.0 + m_box[4] * m_box[4] + m_box[5] * m_box[5]);
        const double Lmax = std::max({a1 / m_p.nx, a2 / m_p.ny, a3 / m_p.nz});
        const double h = Lmax > 0 ? Lmax : 1.0;
        const double K = (kind == math::SplitKind::PSWF) ? m_p.kappa : 2.0 * m_p.kappa * 3.7;
        const double sigma = math::PI / (h * K);
        cw = math::default_window_c(P, sigma);
        }
    m_window = *cached_window(P, m_D, cw);
    // largest |w_x w_y w_z| (bounds the fixed-point contributions of one particle to one cell)
    double wmax = 0.0;
    for (unsigned int k = 0; k < P; ++k)
        for (int i = 0; i <= 512; ++i)
            wmax = std::max(wmax, std::fabs(m_window.eval(k, double(i) / 512.0)));
    m_w3max = float(wmax * wmax * wmax * 1.01);
    m_tables_built = true;
    }

void EspSolver::pair_table(const double** host, const double** device, unsigned int* n_segs)
    {
    if (m_r_lo.empty() || m_need_init)
        build_host_tables();
    const unsigned int n = unsigned(m_r_lo.size());
    m_pair_rec.assign(size_t(n) * 12, 0.0);
    for (unsigned int i = 0; i < n; ++i)
        {
        m_pair_rec[size_t(i) * 12 + 0] = m_r_lo[i];
        m_pair_rec[size_t(i) * 12 + 1] = m_dr_inv[i];
        for (int c = 0; c < 5; ++c)
            {
            m_pair_rec[size_t(i) * 12 + 2 + c] = m_screen[size_t(i) * 5 + c];
            m_pair_rec[size_t(i) * 12 + 7 + c] = m_dscreen[size_t(i) * 5 + c];
            }
        }
    m_pair_rec_dev.resize(m_pair_rec.size() * sizeof(double));
    check_hip(hipMemcpy(m_pair_rec_dev.get(), m_pair_rec.data(), m_pair_rec.size() * sizeof(double), hipMemcpyHostToDevice),
              "hipMemcpy pair table");
    *host = m_pair_rec.data();
    *device = m_pair_rec_dev.as<double>();
    *n_segs = n;
    }

size_t EspSolver::memory_bytes() const
    {
    size_t b = 0;
    const DeviceBuffer* bufs[] = {&m_split_tab, &m_wcoef, &m_tw[0], &m_tw[1], &m_tw[2], &m_G, &m_Gv, &m_qxd, &m_qwd, &m_w2[0], &m_w2[1],
                                  &m_w2[2], &m_mesh, &m_spec, &m_phi, &m_tile_count, &m_tile_start, &m_key, &m_slot, &m_rec,
                                  &m_sidx, &m_qpart, &m_evpart, &m_qsum, &m_ev_internal, &m_pq, &m_compact, &m_pair_rec_dev, &m_fx};
    for (const DeviceBuffer* d : bufs)
        b += d->bytes();
    return b;
    }

void EspSolver::begin_uploads()
    {
    if (m_stage_ev && m_stage_recorded)
        check_hip(hipEventSynchronize(m_stage_ev), "hipEventSynchronize");
    for (unsigned char* p : m_stage_retired)
        check_hip(hipHostFree(p), "hipHostFree");
    m_stage_retired.clear();
    m_stage_recorded = false;
    m_stage_used = 0;
    }

void EspSolver::end_uploads(hipStream_t s)
    {
    if (!m_stage_ev)
        check_hip(hipEventCreateWithFlags(&m_stage_ev, hipEventDisableTiming), "hipEventCreate");
    check_hip(hipEventRecord(m_stage_ev, s), "hipEventRecord");
    m_stage_recorded = true;
    m_stage_stream = s;
    }

//! Convert to R and copy into `buf` through the staging arena (stream-ordered on s).
template<typename R> void EspSolver::upload_as(DeviceBuffer& buf, const std::vector<double>& v, hipStream_t s)
    {
    const size_t bytes = v.size() * sizeof(R);
    buf.resize(bytes);
    if (bytes == 0)
        return;
    size_t off = (m_stage_used + 255) & ~size_t(255);
    if (off + bytes > m_stage_cap)
        {
        // this set-up's earlier copies may still read the old block: it is freed at the next
        // set-up, after they are done
        if (m_stage)
            m_stage_retired.push_back(m_stage);
        m_stage = nullptr;
        const size_t cap = std::max<size_t>(2 * bytes, size_t(2) << 20);
        check_hip(hipHostMalloc(reinterpret_cast<void**>(&m_stage), cap, hipHostMallocDefault), "hipHostMalloc");
        m_stage_cap = cap;
        off = 0;
        }
    R* dst = reinterpret_cast<R*>(m_stage + off);
    for (size_t i = 0; i < v.size(); ++i)
        dst[i] = R(v[i]);
    m_stage_used = off + bytes;
    check_hip(hipMemcpyAsync(buf.get(), dst, bytes, hipMemcpyHostToDevice, s), "hipMemcpyAsync tables");
    }

//! Box lengths > 0 and finite tilt factors (HOOMD's BoxDim: Lx, Ly, Lz, xy, xz, yz).
static void check_box(const double* box)
    {
    if (!(box[0] > 0 && box[1] > 0 && box[2] > 0) || !std::isfinite(box[0]) || !std::isfinite(box[1]) || !std::isfinite(box[2]))
        throw EspError(STATUS_ARG, "box lengths must be finite and > 0");
    if (!std::isfinite(box[3]) || !std::isfinite(box[4]) || !std::isfinite(box[5]))
        throw EspError(STATUS_ARG, "box tilt factors must be finite");
    }

void EspSolver::set_box_geometry(const double* box)
    {
    for (int a = 0; a < 6; ++a)
        m_box[a] = box[a];
    MeshGeom& g = m_geom;
    g.Lx = m_box[0];
    g.Ly = m_box[1];
    g.Lz = m_box[2];
    g.hx = g.Lx / g.nx;
    g.hy = g.Ly / g.ny;
    g.hz = g.Lz / g.nz;
    g.xy = m_box[3];
    g.xz = m_box[4];
    g.yz = m_box[5];
    g.tilted = (g.xy != 0.0 || g.xz != 0.0 || g.yz != 0.0) ? 1 : 0;
    }

void EspSolver::initialize(const SystemView& sys, hipStream_t s)
    {
    check_box(sys.box);
    begin_uploads();
    for (int a = 0; a < 6; ++a)
        m_box[a] = sys.box[a];
    const bool fp64 = m_p.precision == 0;
    const unsigned int P = m_p.order;

    build_host_tables();

    // mesh geometry and tiles: 8-cell tiles unless the LDS box would exceed 64 KiB or the
    // mesh is small, then 4
    MeshGeom& g = m_geom;
    g.nx = int(m_p.nx);
    g.ny = int(m_p.ny);
    g.nz = int(m_p.nz);
    g.nxh = g.nx / 2 + 1;
    g.sxp = (g.nxh + 15) & ~15;
    // tile edge: the requested one, else the largest of {8, 4, 2} whose kernels fit their LDS
    // box, dropping to 4 on small meshes (fewer than 256 tiles of 8) for parallelism
    auto tiles_of = [&](int t) { return size_t((g.nx + t - 1) / t) * ((g.ny + t - 1) / t) * ((g.nz + t - 1) / t); };
    int T = 0;
    if (m_p.tile)
        {
        T = int(m_p.tile);
        if (!kernel::stencil_tile_ok(T, int(P), fp64, false))
            throw EspError(STATUS_UNSUPPORTED, "the requested tile does not fit the stencil kernels' LDS at this order and precision");
        }
    else if (kernel::mfma_spread_tile(int(P), fp64))
        T = kernel::mfma_spread_tile(int(P), fp64);
    else
        {
        for (int t : {8, 4, 2})
            if (kernel::stencil_tile_ok(t, int(P), fp64, false))
                {
                T = t;
                break;
                }
        if (T == 0)
            throw EspError(STATUS_UNSUPPORTED, "stencil order too large for the LDS tile at this precision");
        if (T == 8 && tiles_of(8) < 256 && kernel::stencil_tile_ok(4, int(P), fp64, false))
            T = 4;
        }
    g.T = T;
    g.tx = (g.nx + T - 1) / T;
    g.ty = (g.ny + T - 1) / T;
    g.tz = (g.nz + T - 1) / T;
    g.B = T + int(P) - 1;
    set_box_geometry(sys.box);
    m_n_tiles = unsigned(g.tx) * g.ty * g.tz;
    if (m_n_tiles > (1u << 24))
        throw EspError(STATUS_ARG, "too many spreading tiles");
    m_plan[0] = make_fft_plan(g.nx);
    m_plan[1] = make_fft_plan(g.ny);
    m_plan[2] = make_fft_plan(g.nz);

    // split table records [r_lo, dr_inv, S c0..c4, dS c0..c4]
    const unsigned int nseg = unsigned(m_r_lo.size());
    std::vector<double> rec(size_t(nseg) * 12);
    for (unsigned int i = 0; i < nseg; ++i)
        {
        rec[size_t(i) * 12 + 0] = m_r_lo[i];
        rec[size_t(i) * 12 + 1] = m_dr_inv[i];
        for (int c = 0; c < 5; ++c)
            {
            rec[size_t(i) * 12 + 2 + c] = m_screen[size_t(i) * 5 + c];
            rec[size_t(i) * 12 + 7 + c] = m_dscreen[size_t(i) * 5 + c];
            }
        }
    if (fp64)
        upload_as<double>(m_split_tab, rec, s);
    else
        upload_as<float>(m_split_tab, rec, s);
    // compact LDS-resident table of S for the pair kernel, unless an external
    // reference-format table was supplied (then the reference lookup is used verbatim)
    if (!m_external_table)
        {
        const math::SplitKind kind = m_p.split == 1 ? math::SplitKind::GAUSSIAN : math::SplitKind::PSWF;
        const auto ct = compact_split(kind, m_p.kappa * m_p.rcut, COMPACT_SEGS, compact_deg());
        if (fp64)
            upload_as<double>(m_compact, *ct, s);
        else
            upload_as<float>(m_compact, *ct, s);
        }
    if (fp64)
        upload_as<double>(m_wcoef, m_window.coef, s);
    else
        upload_as<float>(m_wcoef, m_window.coef, s);

    // twiddles exp(-2 pi i m / n)
    const int dims[3] = {g.nx, g.ny, g.nz};
    for (int a = 0; a < 3; ++a)
        {
        std::vector<double> tw(size_t(dims[a]) * 2);
        for (int m = 0; m < dims[a]; ++m)
            {
            const double ang = -2.0 * math::PI * double(m) / double(dims[a]);
            tw[2 * m] = std::cos(ang);
            tw[2 * m + 1] = std::sin(ang);
            }
        if (fp64)
            upload_as<double>(m_tw[a], tw, s);
        else
            upload_as<float>(m_tw[a], tw, s);
        }

    // device buffers
    const size_t M = size_t(g.nx) * g.ny * g.nz;
    const size_t Mh = size_t(g.sxp) * g.ny * g.nz;  // padded rows
    m_mesh.resize_zeroed_async(M * 8, s);
    m_spec.resize(Mh * 2 * (fp64 ? 8 : 4));
    m_phi.resize(M * (fp64 ? 8 : 4));
    m_G.resize(Mh * 8);
    m_Gv.resize(Mh * 8);
    m_tile_count.resize_zeroed_async((size_t(m_n_tiles) + 1) * 4, s);
    m_tile_start.resize_zeroed_async((size_t(m_n_tiles) + 1) * 4, s);
    m_evpart.resize_zeroed_async(size_t(kernel::fft_z_blocks(g, fp64)) * 8 * 8, s);
    m_qsum.resize_zeroed_async(2 * 8, s);
    m_ev_internal.resize_zeroed_async(8 * 8, s);
    m_fx.resize_zeroed_async(4 * sizeof(float), s);
    m_cap_particles = 0;

    build_influence(s);
    end_uploads(s);
    m_need_init = false;
    ++m_gen;
    }

double EspSolver::effective_k_cut() const
    {
    if (m_p.k_cut > 0.0)
        return m_p.k_cut;
    if (m_p.k_cut < 0.0 || m_p.split == 1)
        return 0.0;
    return 1.2 * m_p.kappa;
    }

void EspSolver::build_influence(hipStream_t s)
    {
    const MeshGeom& g = m_geom;
    const math::SplitKind kind = m_p.split == 1 ? math::SplitKind::GAUSSIAN : math::SplitKind::PSWF;
    // |w(theta)|^2 per axis
    std::vector<double> w2x(g.nxh), w2y(g.ny), w2z(g.nz);
    for (int x = 0; x < g.nxh; ++x)
        {
        const double f = m_window.fourier(2.0 * math::PI * x / g.nx);
        w2x[x] = f * f;
        }
    for (int y = 0; y < g.ny; ++y)
        {
        const int k = (y <= g.ny / 2) ? y : y - g.ny;
        const double f = m_window.fourier(2.0 * math::PI * k / g.ny);
        w2y[y] = f * f;
        }
    for (int z = 0; z < g.nz; ++z)
        {
        const int k = (z <= g.nz / 2) ? z : z - g.nz;
        const double f = m_window.fourier(2.0 * math::PI * k / g.nz);
        w2z[z] = f * f;
        }
    upload_as<double>(m_w2[0], w2x, s);
    upload_as<double>(m_w2[1], w2y, s);
    upload_as<double>(m_w2[2], w2z, s);
    int nq = 0;
    if (kind == math::SplitKind::PSWF)
        {
        // largest |k| of the mesh: the corner modes (+-nx/2, +-ny/2, +-nz/2) of k = 2 pi B^-T m
        double kmax = 0.0;
        for (int sy = -1; sy <= 1; sy += 2)
            for (int sz = -1; sz <= 1; sz += 2)
                {
                const double k0x = math::PI / g.hx, k0y = sy * math::PI / g.hy, k0z = sz * math::PI / g.hz;
                const double kx = k0x, ky = k0y - g.xy * k0x, kz = k0z - g.yz * k0y + (g.xy * g.yz - g.xz) * k0x;
                kmax = std::max(kmax, std::sqrt(kx * kx + ky * ky + kz * kz));
                }
        nq = math::SplitFunction::quadrature_nodes(m_p.kappa * m_p.rcut, kmax * m_p.rcut);
        const auto quad = cached_quadrature(kind, m_p.kappa * m_p.rcut, nq);
        m_qx = quad->first;
        m_qw = quad->second;
        upload_as<double>(m_qxd, m_qx, s);
        upload_as<double>(m_qwd, m_qw, s);
        }
    else
        {
        m_qxd.resize(8);
        m_qwd.resize(8);
        }
    check_hip(kernel::build_influence(g, m_p.split, m_p.kappa, m_p.rcut, effective_k_cut(), m_p.alpha, nq, m_qxd.as<double>(), m_qwd.as<double>(), m_w2[0].as<double>(),
                                      m_w2[1].as<double>(), m_w2[2].as<double>(), m_G.as<double>(), m_Gv.as<double>(), s),
              "build_influence");
    }

void EspSolver::ensure_particles(unsigned int n)
    {
    if (n <= m_cap_particles)
        return;
    const unsigned int cap = std::max(n, m_cap_particles + m_cap_particles / 2);
    const bool fp64 = m_p.precision == 0;
    ++m_gen;
    m_key.resize(size_t(cap) * 4);
    m_slot.resize(size_t(cap) * 4);
    m_rec.resize(size_t(cap) * 4 * (fp64 ? 8 : 4));
    m_sidx.resize(size_t(cap) * sizeof(SortedIdx));
    m_qpart.resize(size_t((cap + 255) / 256 + 1) * 3 * 8);
    m_cap_particles = cap;
    }

kernel::PairArgs EspSolver::pair_args(const SystemView& sys, const OutputView& out) const
    {
    kernel::PairArgs a {};
    a.N = sys.N;
    a.postype = sys.postype;
    a.charge = sys.charge;
    a.Lx = sys.box[0];
    a.Ly = sys.box[1];
    a.Lz = sys.box[2];
    a.xy = sys.box[3];
    a.xz = sys.box[4];
    a.yz = sys.box[5];
    a.tilted = (a.xy != 0.0 || a.xz != 0.0 || a.yz != 0.0) ? 1 : 0;
    a.n_neigh = sys.n_neigh;
    a.nlist = sys.nlist;
    a.head_list = sys.head_list;
    a.n_ex = sys.n_ex;
    a.ex_list = sys.ex_list;
    a.ex_pitch = sys.ex_pitch;
    a.split_tab = m_split_tab.get();
    a.n_segs = unsigned(m_r_lo.size());
    a.compact_tab = m_external_table ? nullptr : m_compact.get();
    a.n_compact = COMPACT_SEGS;
    a.compact_deg = compact_deg();
    a.rc = m_p.rcut;
    a.force = out.force;
    a.virial = out.virial;
    a.virial_pitch = out.virial_pitch;
    a.pq = m_pq.as<double>();
    return a;
    }

template<typename R> void EspSolver::run_pair(const SystemView& sys, const OutputView& out, hipStream_t s, bool concurrent)
    {
    const kernel::PairArgs a = pair_args(sys, out);
    int lanes = m_p.pair_lanes ? m_p.pair_lanes : 8;
    // Beside the FP32 mesh chain the pass is held to 3 blocks (12 waves) per CU, which leaves
    // the spreading and FFT kernels room to run alongside it (measured at 1e6 particles, mesh
    // 160^3, r_c = 6.4: 0.99 ms per step against 1.13 ms at 2 blocks, 1.06 ms at 4 and
    // 1.37 ms with the full grid).  The FP64 kernel (124 VGPRs) fits 4 blocks per CU; a grid of
    // exactly that many keeps its blocks from queueing ahead of the mesh chain's (1e6
    // particles, mesh 180^3: 1.84 ms per step against 1.86 ms at 8 blocks per CU, 1.87 at 3).
    const unsigned int cap = concurrent ? (sizeof(R) == 4 ? 3u : 4u) * cu_count() : 0u;
    check_hip(kernel::pack_positions_charges<R>(a, s), "pack_positions_charges");
    // the "pair" stage of the per-stage timing spans the pair kernel alone
    pair_mark(false, s);
    check_hip(kernel::pair_forces<R>(a, lanes, cap, s), "pair_forces");
    pair_mark(true, s);
    }

unsigned int cu_count()
    {
    static std::mutex mu;
    static std::map<int, unsigned int> counts;
    int dev = 0;
    check_hip(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lock(mu);
    auto it = counts.find(dev);
    if (it != counts.end())
        return it->second;
    int n = 0;
    check_hip(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    const unsigned int c = unsigned(std::max(n, 1));
    counts.emplace(dev, c);
    return c;
    }

//! The mesh chain on stream s.  With a side stream (where the pair pass runs) the final FP64
//! reduction also runs there, after the pair pass and the z pass, so it overlaps the inverse y
//! and x passes; the gather waits for the side stream.
template<typename R>
void EspSolver::run_mesh(const SystemView& sys, const OutputView& out, unsigned int n, hipStream_t s, bool want_ev, hipStream_t side)
    {
    const MeshGeom& g = m_geom;
    const int P = int(m_p.order);
    const int D = int(m_D);
    const unsigned int nqb = (n + 255) / 256;
    mark(1, s);
    check_hip(kernel::bin_particles(n, sys.group_idx, sys.postype, sys.charge, g, P, m_tile_count.as<unsigned int>(),
                                    m_key.as<unsigned int>(), m_slot.as<unsigned int>(), m_qpart.as<double>(), s),
              "bin_particles");
    mark(2, s);
    check_hip(kernel::scan_tiles(m_n_tiles, m_tile_count.as<unsigned int>(), m_tile_start.as<unsigned int>(), int(nqb),
                                 m_qpart.as<double>(), fixed_point_rule(), m_fx.as<float>(), s),
              "scan_tiles");
    mark(3, s);
    check_hip(kernel::scatter_particles<R>(n, sys.group_idx, sys.postype, sys.charge, g, P, m_key.as<unsigned int>(),
                                           m_slot.as<unsigned int>(), m_tile_start.as<unsigned int>(), m_rec.get(),
                                           m_sidx.as<SortedIdx>(), s),
              "scatter_particles");
    mark(4, s);
    check_hip(kernel::spread<R>(g, P, D, m_n_tiles, m_tile_start.as<unsigned int>(), m_rec.get(), m_sidx.as<SortedIdx>(), m_wcoef.get(),
                                m_mesh.get(), m_fx.as<float>(), s),
              "spread");
    mark(5, s);
    check_hip(kernel::fft_forward_x<R>(g, m_plan[0], m_tw[0].get(), m_mesh.get(), m_spec.get(), m_fx.as<float>(), s),
              "fft_forward_x");
    mark(6, s);
    check_hip(kernel::fft_y<R>(g, m_plan[1], m_tw[1].get(), m_spec.get(), -1, s), "fft_y");
    mark(7, s);
    check_hip(kernel::fft_z_convolve<R>(g, m_plan[2], m_tw[2].get(), m_spec.get(), m_G.as<double>(), m_Gv.as<double>(), want_ev ? 1 : 0,
                                        m_evpart.as<double>(), s),
              "fft_z_convolve");
    const double V = g.Lx * g.Ly * g.Lz;
    double* evout = out.ext_energy_virial ? out.ext_energy_virial : m_ev_internal.as<double>();
    auto finalize = [&](hipStream_t fs)
    {
        check_hip(kernel::finalize(int(kernel::fft_z_blocks(g, sizeof(R) == 8)), m_evpart.as<double>(), int(nqb), m_qpart.as<double>(),
                                   0.5 / V, m_self_coeff, m_bg_coeff, V, want_ev ? 1 : 0, evout, m_qsum.as<double>(), fs),
                  "finalize");
    };
    if (side)
        {
        check_hip(hipEventRecord(m_ev_zdone, s), "hipEventRecord");
        check_hip(hipStreamWaitEvent(side, m_ev_zdone, 0), "hipStreamWaitEvent");
        finalize(side);
        check_hip(hipEventRecord(m_ev_join, side), "hipEventRecord");
        }
    mark(8, s);
    check_hip(kernel::fft_y<R>(g, m_plan[1], m_tw[1].get(), m_spec.get(), +1, s), "fft_y inverse");
    mark(9, s);
    check_hip(kernel::fft_inverse_x<R>(g, m_plan[0], m_tw[0].get(), m_spec.get(), 1.0 / V, m_phi.get(), s), "fft_inverse_x");
    mark(10, s);
    if (side)
        check_hip(hipStreamWaitEvent(s, m_ev_join, 0), "hipStreamWaitEvent");
    check_hip(kernel::gather<R>(g, P, D, m_n_tiles, m_tile_start.as<unsigned int>(), m_rec.get(), m_sidx.as<SortedIdx>(), m_wcoef.get(),
                                m_phi.get(), out.force, s),
              "gather");
    mark(11, s);
    if (!side)
        finalize(s);
    }

unsigned int EspSolver::prepare(const SystemView& sys, const OutputView& out, hipStream_t s)
    {
    if (!sys.postype || !sys.charge)
        throw EspError(STATUS_ARG, "postype and charge must be device pointers");
    if (!out.force)
        throw EspError(STATUS_ARG, "force output must be a device pointer");
    if (out.virial && out.virial_pitch < sys.N)
        throw EspError(STATUS_ARG, "virial_pitch must be >= N");
    if (sys.nlist && (!sys.n_neigh || !sys.head_list))
        throw EspError(STATUS_ARG, "neighbour list needs n_neigh, nlist and head_list");
    if (sys.n_ex && (!sys.ex_list || sys.ex_pitch == 0))
        throw EspError(STATUS_ARG, "exclusions need ex_list and ex_pitch");
    if (sys.group_idx && sys.group_size > sys.N)
        throw EspError(STATUS_ARG, "group_size exceeds N");
    bool box_changed = false;
    for (int a = 0; a < 6; ++a)
        box_changed = box_changed || sys.box[a] != m_box[a];
    if (m_need_init || box_changed)
        {
        if (m_need_init)
            initialize(sys, s);
        else
            {
            check_box(sys.box);
            begin_uploads();
            set_box_geometry(sys.box);
            build_influence(s);
            end_uploads(s);
            ++m_gen;
            }
        }
    else if (m_stage_recorded && s != m_stage_stream)
        {
        // the set-up ran on another stream: its copies and clears come first
        check_hip(hipStreamWaitEvent(s, m_stage_ev, 0), "hipStreamWaitEvent");
        }
    const unsigned int n = sys.group_idx ? sys.group_size : sys.N;
    ensure_particles(std::max(n, 1u));
    if (m_pq.bytes() < size_t(std::max(sys.N, 1u)) * 32)
        {
        m_pq.resize(size_t(std::max(sys.N, 1u)) * 32 + m_pq.bytes() / 2);
        ++m_gen;
        }
    return n;
    }

void EspSolver::set_graph(bool on)
    {
    m_graph_on = on;
    if (!on)
        release_graph();
    }

void EspSolver::release_graph()
    {
    if (m_gexec)
        (void)hipGraphExecDestroy(m_gexec);
    m_gexec = nullptr;
    m_gkey.clear();
    }

//! Everything a captured step depends on besides device memory contents.
std::vector<uint64_t> EspSolver::graph_key(const SystemView& sys, const OutputView& out, hipStream_t s) const
    {
    auto u = [](const void* p) { return uint64_t(reinterpret_cast<uintptr_t>(p)); };
    std::vector<uint64_t> k = {m_gen, uint64_t(m_overlap), u(s), sys.N, sys.group_size, u(sys.postype), u(sys.charge),
                               u(sys.group_idx), u(sys.n_neigh), u(sys.nlist), u(sys.head_list), u(sys.n_ex), u(sys.ex_list),
                               sys.ex_pitch, u(out.force), u(out.virial), uint64_t(out.virial_pitch), u(out.ext_energy_virial),
                               out.flags};
    for (int a = 0; a < 6; ++a)
        {
        uint64_t b;
        std::memcpy(&b, &sys.box[a], 8);
        k.push_back(b);
        }
    return k;
    }

void EspSolver::compute(const SystemView& sys, const OutputView& out, hipStream_t s)
    {
    const unsigned int n = prepare(sys, out, s);
    const unsigned int flags = out.flags ? out.flags : COMPUTE_ALL;
    // graphs hold no memset nodes: the mesh-only path (which clears the outputs) stays eager
    if (!m_graph_on || m_profile || !(flags & COMPUTE_PAIR))
        {
        enqueue_step(sys, out, n, s);
        return;
        }
    std::vector<uint64_t> key = graph_key(sys, out, s);
    if (!m_gexec || key != m_gkey)
        {
        release_graph();
        if (!m_gstream)
            check_hip(hipStreamCreateWithFlags(&m_gstream, hipStreamNonBlocking), "hipStreamCreate");
        // side-stream objects are created before the capture begins
        if (m_overlap && (flags & COMPUTE_MESH) && !m_side)
            {
            check_hip(hipStreamCreateWithFlags(&m_side, hipStreamNonBlocking), "hipStreamCreate");
            check_hip(hipEventCreateWithFlags(&m_ev_fork, hipEventDisableTiming), "hipEventCreate");
            check_hip(hipEventCreateWithFlags(&m_ev_join, hipEventDisableTiming), "hipEventCreate");
            check_hip(hipEventCreateWithFlags(&m_ev_zdone, hipEventDisableTiming), "hipEventCreate");
            }
        check_hip(hipStreamBeginCapture(m_gstream, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
        hipGraph_t g = nullptr;
        try
            {
            enqueue_step(sys, out, n, m_gstream);
            }
        catch (...)
            {
            (void)hipStreamEndCapture(m_gstream, &g);
            if (g)
                (void)hipGraphDestroy(g);
            throw;
            }
        check_hip(hipStreamEndCapture(m_gstream, &g), "hipStreamEndCapture");
        const hipError_t e = hipGraphInstantiate(&m_gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        check_hip(e, "hipGraphInstantiate");
        m_gkey = std::move(key);
        ++m_gcaptures;
        }
    check_hip(hipGraphLaunch(m_gexec, s), "hipGraphLaunch");
    ++m_greplays;
    }

void EspSolver::enqueue_step(const SystemView& sys, const OutputView& out, unsigned int n, hipStream_t s)
    {
    const bool fp64 = m_p.precision == 0;
    const unsigned int flags = out.flags ? out.flags : COMPUTE_ALL;
    mark(0, s);
    // the pair correction runs on a side stream, overlapping binning, spreading and the FFTs;
    // the gather (which adds into the same force array) waits for it.  (Forking it later in
    // the chain measured no faster.)
    const bool mesh = (flags & COMPUTE_MESH) != 0;
    const bool side = (flags & COMPUTE_PAIR) && mesh && m_overlap;
    if (side && !m_side)
        {
        check_hip(hipStreamCreateWithFlags(&m_side, hipStreamNonBlocking), "hipStreamCreate");
        check_hip(hipEventCreateWithFlags(&m_ev_fork, hipEventDisableTiming), "hipEventCreate");
        check_hip(hipEventCreateWithFlags(&m_ev_join, hipEventDisableTiming), "hipEventCreate");
        check_hip(hipEventCreateWithFlags(&m_ev_zdone, hipEventDisableTiming), "hipEventCreate");
        }
    if (flags & COMPUTE_PAIR)
        {
        hipStream_t ps = s;
        if (side)
            {
            check_hip(hipEventRecord(m_ev_fork, s), "hipEventRecord");
            check_hip(hipStreamWaitEvent(m_side, m_ev_fork, 0), "hipStreamWaitEvent");
            ps = m_side;
            }
        if (fp64)
            run_pair<double>(sys, out, ps, side);
        else
            run_pair<float>(sys, out, ps, side);
        // with the side stream, the mesh chain records the join after the final reduction
        }
    else
        {
        check_hip(hipMemsetAsync(out.force, 0, size_t(sys.N) * 32, s), "hipMemsetAsync force");
        if (out.virial)
            for (int c = 0; c < 6; ++c)
                check_hip(hipMemsetAsync(out.virial + size_t(c) * out.virial_pitch, 0, size_t(sys.N) * 8, s), "hipMemsetAsync virial");
        }
    if (mesh)
        {
        const bool want_ev = (flags & COMPUTE_ENERGY_VIRIAL) != 0;
        if (fp64)
            run_mesh<double>(sys, out, n, s, want_ev, side ? m_side : nullptr);
        else
            run_mesh<float>(sys, out, n, s, want_ev, side ? m_side : nullptr);
        }
    else
        for (int b = 1; b <= 11; ++b)
            mark(b, s);
    mark(N_STAGES, s);
    }

void EspSolver::pair_forces_into(const SystemView& sys, const OutputView& out, hipStream_t s, void* scratch, size_t scratch_bytes)
    {
    if (m_need_init || !m_split_tab.get())
        throw EspError(STATUS_STATE, "the context is not initialised: one esp_compute_forces or esp_pair_forces call (which "
                                     "allocates) must follow esp_create / esp_set_params before the no-allocation pair entry point");
    if (sys.N == 0)
        return;
    if (!sys.postype || !sys.charge || !out.force)
        throw EspError(STATUS_ARG, "postype, charge and force must be device pointers");
    if (!sys.nlist || !sys.n_neigh || !sys.head_list)
        throw EspError(STATUS_ARG, "the pair correction needs n_neigh, nlist and head_list");
    if (out.virial && out.virial_pitch < sys.N)
        throw EspError(STATUS_ARG, "virial_pitch must be >= N");
    if (sys.n_ex && (!sys.ex_list || sys.ex_pitch == 0))
        throw EspError(STATUS_ARG, "exclusions need ex_list and ex_pitch");
    check_box(sys.box);
    if (size_t(sys.N) * 32 > scratch_bytes)
        throw EspError(STATUS_RANGE, "N exceeds the particles the no-allocation scratch holds");
    kernel::PairArgs a = pair_args(sys, out);
    a.pq = static_cast<double*>(scratch);
    const int lanes = m_p.pair_lanes ? m_p.pair_lanes : 8;
    if (m_p.precision == 0)
        {
        check_hip(kernel::pack_positions_charges<double>(a, s), "pack_positions_charges");
        check_hip(kernel::pair_forces<double>(a, lanes, 0u, s), "pair_forces");
        }
    else
        {
        check_hip(kernel::pack_positions_charges<float>(a, s), "pack_positions_charges");
        check_hip(kernel::pair_forces<float>(a, lanes, 0u, s), "pair_forces");
        }
    }

void EspSolver::describe(const SystemView& sys, const OutputView& out, hipStream_t s, kernel::BatchSys& d)
    {
    if (!kernel::batch_order_supported(int(m_p.order)))
        throw EspError(STATUS_UNSUPPORTED, "batched solves support stencil orders 4, 6, 8, 10 and 12");
    const unsigned int n = prepare(sys, out, s);
    if (!kernel::stencil_tile_ok(m_geom.T, int(m_p.order), m_p.precision == 0, true))
        throw EspError(STATUS_UNSUPPORTED, "the tile of this system does not fit the batched stencil kernels' LDS");
    const bool fp64 = m_p.precision == 0;
    const MeshGeom& g = m_geom;
    d = kernel::BatchSys {};
    d.g = g;
    for (int a = 0; a < 3; ++a)
        d.plan[a] = m_plan[a];
    d.P = int(m_p.order);
    d.D = int(m_D);
    d.n = n;
    d.n_tiles = m_n_tiles;
    d.flags = out.flags ? out.flags : COMPUTE_ALL;
    d.want_ev = (d.flags & COMPUTE_ENERGY_VIRIAL) ? 1 : 0;
    d.ppb_x = kernel::fft_x_pairs_per_block(g.nx, fp64);
    d.ncol_y = kernel::fft_cols_per_block(g.ny, g.nxh, fp64);
    d.ncol_z = kernel::fft_cols_per_block(g.nz, g.nxh, fp64);
    d.n_qblk = (n + kernel::BATCH_CHUNK - 1) / kernel::BATCH_CHUNK;
    d.n_zblk = kernel::fft_z_blocks(g, fp64);
    d.V = g.Lx * g.Ly * g.Lz;
    d.inv_2V = 0.5 / d.V;
    d.self_coeff = m_self_coeff;
    d.bg_coeff = m_bg_coeff;
    d.group = sys.group_idx;
    d.postype = sys.postype;
    d.charge = sys.charge;
    d.force = out.force;
    d.wcoef = m_wcoef.get();
    for (int a = 0; a < 3; ++a)
        d.tw[a] = m_tw[a].get();
    d.G = m_G.as<double>();
    d.Gv = m_Gv.as<double>();
    d.mesh = m_mesh.get();
    d.spec = m_spec.get();
    d.phi = m_phi.get();
    d.tile_count = m_tile_count.as<unsigned int>();
    d.tile_start = m_tile_start.as<unsigned int>();
    d.key = m_key.as<unsigned int>();
    d.slot = m_slot.as<unsigned int>();
    d.rec = m_rec.get();
    d.sidx = m_sidx.as<SortedIdx>();
    d.qpart = m_qpart.as<double>();
    d.fx = m_fx.as<float>();
    d.fxr = fixed_point_rule();
    d.evpart = m_evpart.as<double>();
    d.qsum = m_qsum.as<double>();
    d.ev_out = out.ext_energy_virial ? out.ext_energy_virial : m_ev_internal.as<double>();
    d.pair = pair_args(sys, out);
    }

void EspSolver::fixed_point_scale(float& scale, int& lowered)
    {
    float h[4] = {0, 0, 0, 0};
    if (m_fx.get())
        {
        check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
        check_hip(hipMemcpy(h, m_fx.get(), sizeof(h), hipMemcpyDeviceToHost), "hipMemcpy fx");
        }
    scale = h[0];
    lowered = h[2] != 0.0f ? 1 : 0;
    }

void EspSolver::charge_sums(double& q, double& q2)
    {
    double h[2] = {0, 0};
    if (m_qsum.get())
        {
        check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
        check_hip(hipMemcpy(h, m_qsum.get(), 16, hipMemcpyDeviceToHost), "hipMemcpy qsum");
        }
    q = h[0];
    q2 = h[1];
    }

    } // namespace esp
