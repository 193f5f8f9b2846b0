// Batched ESP solves (see esp_batch.h).
#include "esp_batch.h"

#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <set>

namespace esp
{

namespace
{
size_t fft_lds(int n, int cols, size_t complex_bytes)
    {
    return (size_t(n) + 2 * size_t(cols) * (size_t(n) + 1)) * complex_bytes;
    }

//! The fused xy-plane FFT passes for small meshes (on unless ESP_BATCH_FUSED_XY=0; both forms
//! give the same bits).
bool fused_xy_enabled()
    {
    static const bool on = [] {
        const char* e = getenv("ESP_BATCH_FUSED_XY");
        return !(e && e[0] == '0');
    }();
    return on;
    }
} // namespace

EspBatch::EspBatch() = default;

EspBatch::~EspBatch()
    {
    if (m_done)
        (void)hipEventSynchronize(m_done);
    release_graph();
    if (m_staging)
        (void)hipHostFree(m_staging);
    if (m_items_staging)
        (void)hipHostFree(m_items_staging);
    for (hipEvent_t e : {m_fork, m_join, m_done, m_upload_done})
        if (e)
            (void)hipEventDestroy(e);
    for (hipStream_t st : {m_cap, m_side})
        if (st)
            (void)hipStreamDestroy(st);
    }

void EspBatch::release_graph()
    {
    if (m_exec)
        (void)hipGraphExecDestroy(m_exec);
    if (m_graph)
        (void)hipGraphDestroy(m_graph);
    m_exec = nullptr;
    m_graph = nullptr;
    }

void EspBatch::set_options(bool use_graph, unsigned int max_grid)
    {
    if (m_done)
        check_hip(hipEventSynchronize(m_done), "hipEventSynchronize");
    m_use_graph = use_graph;
    m_max_grid = max_grid;
    release_graph();
    m_shape.clear();  // re-plan with the new grid sizes
    }

void EspBatch::build_plan(bool fp64, hipStream_t s)
    {
    using kernel::BatchItem;
    const unsigned int n = unsigned(m_desc.size());
    const size_t cplx = fp64 ? 16 : 8;
    std::vector<BatchItem> pack, pair, bin, scan, tiles, fx, fy, fz, fxy;
    m_lds_stencil = m_lds_fx = m_lds_fy = m_lds_fz = m_lds_fxy = 0;
    m_lds_pair = 0;
    m_any_mesh = false;
    m_any_tilted = false;
    const unsigned int per_chunk = 256u / unsigned(kernel::BATCH_PAIR_LANES);
    for (unsigned int s = 0; s < n; ++s)
        {
        const kernel::BatchSys& d = m_desc[s];
        m_any_tilted = m_any_tilted || d.pair.tilted;
        for (unsigned int c = 0; c < (d.pair.N + per_chunk - 1) / per_chunk; ++c)
            pair.push_back({s, c});
        for (unsigned int c = 0; c < (d.pair.N + 255u) / 256u; ++c)
            pack.push_back({s, c});
        if (d.pair.compact_tab)
            m_lds_pair = std::max(m_lds_pair, size_t(d.pair.n_compact) * (d.pair.compact_deg + 1) * (fp64 ? 8 : 4));
        if (!(d.flags & COMPUTE_MESH))
            continue;
        m_any_mesh = true;
        for (unsigned int c = 0; c < d.n_qblk; ++c)
            bin.push_back({s, c});
        scan.push_back({s, 0});
        for (unsigned int t = 0; t < d.n_tiles; ++t)
            tiles.push_back({s, t});
        const MeshGeom& g = d.g;
        if (fused_xy_enabled() && kernel::fft_xy_fused(g, fp64))
            {
            // small planes: the x and y passes fused, one workgroup per plane
            for (int z = 0; z < g.nz; ++z)
                fxy.push_back({s, unsigned(z)});
            m_lds_fxy = std::max(m_lds_fxy, kernel::fft_xy_lds(g, fp64));
            }
        else
            {
            const unsigned int npairs = unsigned((g.ny + 1) / 2) * unsigned(g.nz);
            for (unsigned int b = 0; b < (npairs + d.ppb_x - 1) / unsigned(d.ppb_x); ++b)
                fx.push_back({s, b});
            const unsigned int nby = unsigned((g.nxh + d.ncol_y - 1) / d.ncol_y) * unsigned(g.nz);
            for (unsigned int b = 0; b < nby; ++b)
                fy.push_back({s, b});
            m_lds_fx = std::max(m_lds_fx, fft_lds(g.nx, d.ppb_x, cplx));
            m_lds_fy = std::max(m_lds_fy, fft_lds(g.ny, d.ncol_y, cplx));
            }
        for (unsigned int b = 0; b < d.n_zblk; ++b)
            fz.push_back({s, b});
        m_lds_stencil = std::max(m_lds_stencil, kernel::tp_lds_bytes(g, d.P, fp64));
        m_lds_fz = std::max(m_lds_fz, fft_lds(g.nz, d.ncol_z, cplx));
        }
    for (size_t lds : {m_lds_stencil, m_lds_fx, m_lds_fy, m_lds_fz, m_lds_fxy, m_lds_pair})
        if (lds > 64 * 1024)
            throw EspError(STATUS_UNSUPPORTED, "a batched stage needs more than 64 KiB of LDS");

    std::vector<BatchItem> all;
    all.reserve(pack.size() + pair.size() + bin.size() + scan.size() + tiles.size() + fx.size() + fy.size() + fz.size() +
                fxy.size());
    auto place = [&](const std::vector<BatchItem>& v, size_t& off) {
        off = all.size();
        all.insert(all.end(), v.begin(), v.end());
    };
    size_t o_pack, o_pair, o_bin, o_scan, o_tiles, o_fx, o_fy, o_fz, o_fxy;
    place(pack, o_pack);
    place(pair, o_pair);
    place(bin, o_bin);
    place(scan, o_scan);
    place(tiles, o_tiles);
    place(fx, o_fx);
    place(fy, o_fy);
    place(fz, o_fz);
    place(fxy, o_fxy);
    m_items_dev.resize(std::max<size_t>(all.size(), 1) * sizeof(BatchItem));
    if (!all.empty())
        {
        // stream-ordered from pinned staging (a blocking hipMemcpy queued behind the work already
        // on the device); the caller has waited for the launches that read the previous lists
        if (m_items_staging_cap < all.size())
            {
            if (m_items_staging)
                check_hip(hipHostFree(m_items_staging), "hipHostFree");
            m_items_staging = nullptr;
            m_items_staging_cap = 0;
            const size_t cap = all.size() + all.size() / 2;
            check_hip(hipHostMalloc(reinterpret_cast<void**>(&m_items_staging), cap * sizeof(BatchItem), hipHostMallocDefault),
                      "hipHostMalloc");
            m_items_staging_cap = cap;
            }
        std::memcpy(m_items_staging, all.data(), all.size() * sizeof(BatchItem));
        check_hip(hipMemcpyAsync(m_items_dev.get(), m_items_staging, all.size() * sizeof(BatchItem), hipMemcpyHostToDevice, s),
                  "hipMemcpyAsync items");
        }
    const BatchItem* base = m_items_dev.as<BatchItem>();
    auto list = [&](size_t off, size_t count, unsigned int cap) {
        kernel::BatchList l;
        l.items = base + off;
        l.n = unsigned(count);
        l.grid = unsigned(cap ? std::min<size_t>(count, cap) : count);
        if (l.grid == 0)
            l.grid = 1;
        return l;
    };
    m_l.pack = list(o_pack, pack.size(), m_max_grid);
    // The pair branch runs beside the mesh stages.  In FP32 its grid is held to 2 blocks per
    // CU, which leaves the mesh kernels room (512 mixed systems of 500-50,000 particles, 12^3
    // tiles: 8.5-8.8 ms per batch step against 8.7 ms at 1.5 blocks per CU, 9.3 at 1, 9.3-9.7
    // at 3-4 and 10.6 with 2048 blocks).  Environment ESP_BATCH_PAIR_GRID overrides the cap.
    static long env_pair_grid = -1;
    if (env_pair_grid < 0)
        {
        const char* e = getenv("ESP_BATCH_PAIR_GRID");
        env_pair_grid = e ? std::max(8L, atol(e)) : 0L;
        }
    unsigned int pair_cap = m_max_grid;
    if (env_pair_grid)
        pair_cap = unsigned(env_pair_grid);
    else if (!fp64)
        pair_cap = pair_cap ? std::min(pair_cap, 2u * cu_count()) : 2u * cu_count();
    m_l.pair = list(o_pair, pair.size(), pair_cap);
    m_l.bin = list(o_bin, bin.size(), m_max_grid);
    m_l.scatter = m_l.bin;
    m_l.scan = list(o_scan, scan.size(), 0);
    m_l.fin = m_l.scan;
    m_l.tiles = list(o_tiles, tiles.size(), m_max_grid);
    m_l.fft_x = list(o_fx, fx.size(), m_max_grid);
    m_l.fft_y = list(o_fy, fy.size(), m_max_grid);
    m_l.fft_z = list(o_fz, fz.size(), m_max_grid);
    m_l.fft_xy = list(o_fxy, fxy.size(), m_max_grid);
    }

template<typename R> void EspBatch::enqueue_stages(hipStream_t st)
    {
    const kernel::BatchSys* d = m_desc_dev.as<kernel::BatchSys>();
    // the pair correction runs on a branch; the gather (which adds into the same force
    // arrays) waits for it
    check_hip(kernel::batch_pack<R>(d, m_l.pack, st), "batch_pack");
    check_hip(hipEventRecord(m_fork, st), "hipEventRecord");
    check_hip(hipStreamWaitEvent(m_side, m_fork, 0), "hipStreamWaitEvent");
    check_hip(kernel::batch_pair<R>(d, m_l.pair, m_lds_pair, m_any_tilted, m_side), "batch_pair");
    check_hip(hipEventRecord(m_join, m_side), "hipEventRecord");
    if (m_any_mesh)
        {
        check_hip(kernel::batch_bin(d, m_l.bin, st), "batch_bin");
        check_hip(kernel::batch_scan(d, m_l.scan, st), "batch_scan");
        check_hip(kernel::batch_scatter<R>(d, m_l.scatter, st), "batch_scatter");
        check_hip(kernel::batch_spread<R>(d, m_l.tiles, m_lds_stencil, st), "batch_spread");
        check_hip(kernel::batch_fft_xy_forward<R>(d, m_l.fft_xy, m_lds_fxy, st), "batch_fft_xy_forward");
        check_hip(kernel::batch_fft_x_r2c<R>(d, m_l.fft_x, m_lds_fx, st), "batch_fft_x_r2c");
        check_hip(kernel::batch_fft_y<R>(d, m_l.fft_y, -1, m_lds_fy, st), "batch_fft_y");
        check_hip(kernel::batch_fft_z_conv<R>(d, m_l.fft_z, m_lds_fz, st), "batch_fft_z_conv");
        check_hip(kernel::batch_fft_y<R>(d, m_l.fft_y, +1, m_lds_fy, st), "batch_fft_y inverse");
        check_hip(kernel::batch_fft_x_c2r<R>(d, m_l.fft_x, m_lds_fx, st), "batch_fft_x_c2r");
        check_hip(kernel::batch_fft_xy_inverse<R>(d, m_l.fft_xy, m_lds_fxy, st), "batch_fft_xy_inverse");
        }
    check_hip(hipStreamWaitEvent(st, m_join, 0), "hipStreamWaitEvent");
    if (m_any_mesh)
        {
        check_hip(kernel::batch_gather<R>(d, m_l.tiles, m_lds_stencil, st), "batch_gather");
        check_hip(kernel::batch_finalize(d, m_l.fin, st), "batch_finalize");
        }
    }

void EspBatch::enqueue(hipStream_t s)
    {
    // Every shape is captured, including one that lasts a single call: replayed, the pair
    // branch overlaps the mesh chain (512 mixed systems: 7.3 ms per step), launched eagerly on
    // two streams it does not (8.7-8.9 ms), which outweighs a capture's ~0.3 ms of host time.
    if (!m_use_graph)
        {
        if (m_fp64)
            enqueue_stages<double>(s);
        else
            enqueue_stages<float>(s);
        ++m_launches;
        return;
        }
    if (!m_exec)
        {
        check_hip(hipStreamBeginCapture(m_cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try
            {
            if (m_fp64)
                enqueue_stages<double>(m_cap);
            else
                enqueue_stages<float>(m_cap);
            }
        catch (...)
            {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(m_cap, &g);
            if (g)
                (void)hipGraphDestroy(g);
            throw;
            }
        check_hip(hipStreamEndCapture(m_cap, &m_graph), "hipStreamEndCapture");
        check_hip(hipGraphInstantiate(&m_exec, m_graph, nullptr, nullptr, 0), "hipGraphInstantiate");
        ++m_captures;
        }
    check_hip(hipGraphLaunch(m_exec, s), "hipGraphLaunch");
    ++m_replays;
    }

void EspBatch::compute(unsigned int n, EspSolver* const* solvers, const SystemView* sys, const OutputView* out, hipStream_t s)
    {
    if (n == 0)
        return;
    if (!solvers || !sys || !out)
        throw EspError(STATUS_ARG, "null batch arrays");
    std::set<const EspSolver*> seen;
    for (unsigned int i = 0; i < n; ++i)
        {
        if (!solvers[i])
            throw EspError(STATUS_ARG, "null solver in batch");
        if (!seen.insert(solvers[i]).second)
            throw EspError(STATUS_ARG, "a solver appears twice in one batch (its scratch would be shared)");
        }
    const bool fp64 = solvers[0]->fp64();
    for (unsigned int i = 1; i < n; ++i)
        if (solvers[i]->fp64() != fp64)
            throw EspError(STATUS_ARG, "all systems of a batch must use the same precision");
    if (!m_cap)
        {
        check_hip(hipStreamCreateWithFlags(&m_cap, hipStreamNonBlocking), "hipStreamCreate");
        check_hip(hipStreamCreateWithFlags(&m_side, hipStreamNonBlocking), "hipStreamCreate");
        check_hip(hipEventCreateWithFlags(&m_fork, hipEventDisableTiming), "hipEventCreate");
        check_hip(hipEventCreateWithFlags(&m_join, hipEventDisableTiming), "hipEventCreate");
        check_hip(hipEventCreateWithFlags(&m_done, hipEventDisableTiming), "hipEventCreate");
        check_hip(hipEventCreateWithFlags(&m_upload_done, hipEventDisableTiming), "hipEventCreate");
        }

    m_desc.resize(n);
    for (unsigned int i = 0; i < n; ++i)
        solvers[i]->describe(sys[i], out[i], s, m_desc[i]);

    // shape key: everything that fixes the work lists and launch sizes
    std::vector<uint64_t> shape;
    shape.reserve(size_t(n) * 13 + 2);
    shape.push_back(n);
    shape.push_back(fp64 ? 1 : 0);
    for (const kernel::BatchSys& d : m_desc)
        {
        const uint64_t v[] = {d.n, d.pair.N, d.n_tiles, uint64_t(d.g.nx), uint64_t(d.g.ny), uint64_t(d.g.nz), uint64_t(d.P),
                              uint64_t(d.ppb_x), uint64_t(d.ncol_y), uint64_t(d.ncol_z), d.flags, d.pair.compact_tab ? 1u : 0u,
                              uint64_t(d.pair.tilted)};
        shape.insert(shape.end(), std::begin(v), std::end(v));
        }
    const bool reshape = shape != m_shape;
    const size_t bytes = size_t(n) * sizeof(kernel::BatchSys);
    const bool reupload = reshape || m_uploaded.size() != n || std::memcmp(m_uploaded.data(), m_desc.data(), bytes) != 0;
    if (reshape || reupload)
        check_hip(hipEventSynchronize(m_done), "hipEventSynchronize");  // earlier launches may still read the tables
    if (reshape)
        {
        release_graph();
        m_fp64 = fp64;
        build_plan(fp64, s);
        m_shape = shape;
        }
    if (reupload)
        {
        if (m_staging_cap < n)
            {
            if (m_staging)
                check_hip(hipHostFree(m_staging), "hipHostFree");
            m_staging = nullptr;
            check_hip(hipHostMalloc(reinterpret_cast<void**>(&m_staging), bytes, hipHostMallocDefault), "hipHostMalloc");
            m_staging_cap = n;
            }
        const bool grew = m_desc_dev.bytes() < bytes;
        m_desc_dev.resize(bytes);
        if (grew)
            release_graph();  // the captured launches point at the old table
        std::memcpy(m_staging, m_desc.data(), bytes);
        check_hip(hipMemcpyAsync(m_desc_dev.get(), m_staging, bytes, hipMemcpyHostToDevice, s), "hipMemcpyAsync descriptors");
        check_hip(hipEventRecord(m_upload_done, s), "hipEventRecord");
        m_uploaded = m_desc;
        }
    enqueue(s);
    check_hip(hipEventRecord(m_done, s), "hipEventRecord");
    }

    } // namespace esp
