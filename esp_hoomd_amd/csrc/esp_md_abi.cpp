// C ABI of the MD support: neighbour-list objects, WCA forces, velocity-Verlet / Langevin
// integration and energy reductions.
#include "../../include/esp_mi355x.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "esp_kernels.h"
#include "esp_md.h"
#include "esp_solver.h"

struct esp_nlist
    {
    double r_cut = 0.0;
    double r_buff = 0.0;
    double r_inner = 0.0;      // 0: no inner list
    unsigned int nmax = 32;
    unsigned int nmax_in = 16;
    bool inner_built = false;
    unsigned int N = 0;
    unsigned int n_cells = 0;
    double box[3] = {0, 0, 0};
    bool built = false;
    unsigned long long builds = 0;
    esp::DeviceBuffer n_neigh, nlist, head, cell, slot, count, start, unsorted, members, spos, overflow, last_pos, disp;
    esp::DeviceBuffer n_inner, inner, head_inner;
    unsigned int* h_over = nullptr;  // pinned: overflow counts of the last build (full, inner)
    esp::md::CellGeom cg {};
    bool want_inner = false;
    bool fixed_capacity = false;     // row capacities set by esp_nlist_set_capacity (not from the density)
    std::string last_error;
    ~esp_nlist()
        {
        if (h_over)
            (void)hipHostFree(h_over);
        }
    };

namespace
{
thread_local std::string g_md_error;

int fail(esp_nlist* nl, int code, const std::string& msg)
    {
    if (nl)
        nl->last_error = msg;
    g_md_error = msg;
    return code;
    }

template<class Fn> int guard(esp_nlist* nl, Fn fn)
    {
    try
        {
        fn();
        return ESP_OK;
        }
    catch (const esp::EspError& e)
        {
        return fail(nl, e.code(), e.what());
        }
    catch (const std::exception& e)
        {
        return fail(nl, ESP_ERR_STATE, e.what());
        }
    }

//! Size the list's buffers for this system and fix the cell geometry (host work only).
void build_prepare(esp_nlist* nl, const esp_system* sys)
    {
    using esp::check_hip;
    const unsigned int N = sys->N;
    if (sys->box[3] != 0.0 || sys->box[4] != 0.0 || sys->box[5] != 0.0)
        throw esp::EspError(ESP_ERR_UNSUPPORTED, "triclinic boxes are not supported");
    const double r_list = nl->r_cut + nl->r_buff;
    const double Lmin = std::min({sys->box[0], sys->box[1], sys->box[2]});
    if (!(r_list > 0.0) || r_list >= 0.5 * Lmin)
        throw esp::EspError(ESP_ERR_ARG, "r_cut + r_buff must be positive and below half the box length");
    esp::md::CellGeom& cg = nl->cg;
    cg.Lx = sys->box[0];
    cg.Ly = sys->box[1];
    cg.Lz = sys->box[2];
    // Cells of at least r_list / 2 searched two deep when every axis has five or more of them
    // (a search volume of (5/2)^3 r_list^3 instead of 27 r_list^3: 1e6 particles at
    // r_list = 6.67 test 460 candidates each instead of 800), else cells of at least r_list
    // searched one deep.
    const int hx = int(std::floor(cg.Lx / (0.5 * r_list))), hy = int(std::floor(cg.Ly / (0.5 * r_list))),
              hz = int(std::floor(cg.Lz / (0.5 * r_list)));
    if (std::min({hx, hy, hz}) >= 5 && size_t(hx) * hy * hz <= (1u << 22))
        {
        cg.nx = hx;
        cg.ny = hy;
        cg.nz = hz;
        cg.reach = 2;
        }
    else
        {
        cg.nx = std::max(1, int(std::floor(cg.Lx / r_list)));
        cg.ny = std::max(1, int(std::floor(cg.Ly / r_list)));
        cg.nz = std::max(1, int(std::floor(cg.Lz / r_list)));
        cg.reach = 1;
        // keep the cell count bounded for the single-workgroup scan
        while (size_t(cg.nx) * cg.ny * cg.nz > (1u << 22))
            {
            cg.nx = std::max(1, cg.nx / 2);
            cg.ny = std::max(1, cg.ny / 2);
            cg.nz = std::max(1, cg.nz / 2);
            }
        }
    const unsigned int nc = unsigned(cg.nx * cg.ny * cg.nz);
    if (N > nl->N || nc != nl->n_cells || !nl->n_neigh.get())
        {
        nl->n_neigh.resize(size_t(std::max(N, 1u)) * 4);
        nl->head.resize(size_t(std::max(N, 1u)) * 8);
        nl->cell.resize(size_t(std::max(N, 1u)) * 4);
        nl->slot.resize(size_t(std::max(N, 1u)) * 4);
        nl->members.resize(size_t(std::max(N, 1u)) * 4);
        nl->unsorted.resize(size_t(std::max(N, 1u)) * 4);
        nl->spos.resize(size_t(std::max(N, 1u)) * 32);
        nl->last_pos.resize(size_t(std::max(N, 1u)) * 32);
        nl->count.resize(size_t(nc + 1) * 4, true);
        nl->start.resize(size_t(nc + 1) * 4, true);
        nl->overflow.resize(8, true);
        nl->disp.resize(8, true);
        nl->N = N;
        nl->n_cells = nc;
        }
    // Row capacities from the mean density (1.5x the mean count plus 32, in multiples of 8), so
    // a first build rarely overflows and has to run again; they only ever grow.
    const double rho = double(N) / (cg.Lx * cg.Ly * cg.Lz);
    const double sphere = 4.18879020478639 * r_list * r_list * r_list;
    const unsigned int est = (unsigned(1.5 * rho * sphere) + 32u + 7u) & ~7u;
    if (!nl->fixed_capacity)
        nl->nmax = std::max(nl->nmax, std::min(est, std::max(N, 8u)));
    nl->want_inner = nl->r_inner > 0.0 && nl->r_inner < nl->r_cut;
    if (nl->want_inner && !nl->fixed_capacity)
        {
        const double ri = nl->r_inner + nl->r_buff;
        const unsigned int est_in = (unsigned(1.5 * rho * 4.18879020478639 * ri * ri * ri) + 16u + 7u) & ~7u;
        nl->nmax_in = std::max(nl->nmax_in, std::min(est_in, std::max(N, 8u)));
        }
    if (nl->want_inner)
        {
        nl->n_inner.resize(size_t(std::max(N, 1u)) * 4);
        nl->head_inner.resize(size_t(std::max(N, 1u)) * 8);
        }
    if (!nl->h_over)
        check_hip(hipHostMalloc(reinterpret_cast<void**>(&nl->h_over), 2 * sizeof(unsigned int), hipHostMallocDefault), "hipHostMalloc");
    nl->inner_built = false;
    }

//! Enqueue one build attempt at the current capacities; its overflow counts land in h_over
//! when the stream reaches them.
void build_enqueue(esp_nlist* nl, const esp_system* sys, hipStream_t s)
    {
    using esp::check_hip;
    const unsigned int N = sys->N;
    const unsigned int nc = nl->n_cells;
    const double r_list = nl->r_cut + nl->r_buff;
    nl->nlist.resize(size_t(std::max(N, 1u)) * nl->nmax * 4);
    esp::md::InnerList in {};
    if (nl->want_inner)
        {
        nl->inner.resize(size_t(std::max(N, 1u)) * nl->nmax_in * 4);
        in.r2 = (nl->r_inner + nl->r_buff) * (nl->r_inner + nl->r_buff);
        in.nmax = nl->nmax_in;
        in.n = nl->n_inner.as<unsigned int>();
        in.list = nl->inner.as<unsigned int>();
        in.head = nl->head_inner.as<uint64_t>();
        }
    check_hip(hipMemsetAsync(nl->count.get(), 0, size_t(nc + 1) * 4, s), "hipMemsetAsync");
    check_hip(hipMemsetAsync(nl->overflow.get(), 0, 8, s), "hipMemsetAsync");
    check_hip(esp::md::cell_count(N, sys->postype, nl->cg, nl->count.as<unsigned int>(), nl->cell.as<unsigned int>(),
                                  nl->slot.as<unsigned int>(), s),
              "cell_count");
    check_hip(esp::kernel::scan_tiles(nc, nl->count.as<unsigned int>(), nl->start.as<unsigned int>(), 0, nullptr,
                                      esp::FixedPointRule {}, nullptr, s),
              "scan");
    check_hip(esp::md::cell_fill(N, sys->postype, nl->cg, nl->cell.as<unsigned int>(), nl->slot.as<unsigned int>(),
                                 nl->start.as<unsigned int>(), nl->unsorted.as<unsigned int>(), nl->members.as<unsigned int>(),
                                 nl->spos.as<double>(), s),
              "cell_fill");
    check_hip(esp::md::build_nlist(N, nl->spos.as<double>(), nl->cg, nl->cell.as<unsigned int>(), nl->start.as<unsigned int>(),
                                   nl->members.as<unsigned int>(), r_list,
                                   sys->n_ex, sys->ex_list, sys->ex_pitch, nl->nmax, nl->n_neigh.as<unsigned int>(),
                                   nl->nlist.as<unsigned int>(), nl->head.as<uint64_t>(), nl->overflow.as<unsigned int>(), in, s),
              "build_nlist");
    check_hip(hipMemcpyAsync(nl->h_over, nl->overflow.get(), 8, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    }

//! After the stream has passed build_enqueue: accept the list (true) or raise the capacities
//! the attempt overflowed (false: enqueue again).
bool build_finish(esp_nlist* nl, const esp_system* sys, hipStream_t s)
    {
    const unsigned int over0 = nl->h_over[0], over1 = nl->h_over[1];
    if (over0 == 0 && over1 == 0)
        {
        nl->inner_built = nl->want_inner;
        esp::check_hip(esp::md::copy4(sys->N, sys->postype, nl->last_pos.as<double>(), s), "copy4");
        nl->box[0] = sys->box[0];
        nl->box[1] = sys->box[1];
        nl->box[2] = sys->box[2];
        nl->built = true;
        nl->builds += 1;
        return true;
        }
    if (over0)
        nl->nmax = ((over0 + over0 / 4) + 7) & ~7u;
    if (over1)
        nl->nmax_in = ((over1 + over1 / 4) + 7) & ~7u;
    return false;
    }

//! Build now: enqueue, wait, and retry with larger capacities while the list overflows.
void do_build(esp_nlist* nl, const esp_system* sys, hipStream_t s)
    {
    build_prepare(nl, sys);
    for (int attempt = 0; attempt < 4; ++attempt)
        {
        build_enqueue(nl, sys, s);
        esp::check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
        if (build_finish(nl, sys, s))
            return;
        }
    throw esp::EspError(ESP_ERR_STATE, "neighbour list did not converge");
    }
} // namespace

extern "C" {

int esp_nlist_create(esp_nlist** nl, double r_cut, double r_buff)
    {
    if (!nl || !(r_cut > 0.0) || r_buff < 0.0)
        return fail(nullptr, ESP_ERR_ARG, "esp_nlist_create: r_cut must be > 0 and r_buff >= 0");
    *nl = new esp_nlist();
    (*nl)->r_cut = r_cut;
    (*nl)->r_buff = r_buff;
    return ESP_OK;
    }

int esp_nlist_destroy(esp_nlist* nl)
    {
    delete nl;
    return ESP_OK;
    }

} // extern "C"

// ---- batched MD: many systems advanced by one launch per operation -----------------------
struct esp_md_batch
    {
    std::vector<esp_system> sys;
    std::vector<esp_nlist*> nls;
    std::vector<esp::md::MDSys> desc;
    std::vector<esp::md::MDItem> items;
    esp::DeviceBuffer d_disp_pred;  // predicted displacements of the last esp_md_batch_predict
    esp::DeviceBuffer d_desc, d_items, d_disp;
    std::vector<unsigned long long> h_disp;
    unsigned long long* h_pred = nullptr;  // pinned copy of d_disp_pred
    unsigned int pred_cap = 0;
    std::vector<unsigned int> pending;      // members whose asynchronous rebuild awaits its check
    void* h_stage = nullptr;                // pinned staging of descriptor and item uploads
    size_t stage_cap = 0;
    esp::DeviceBuffer d_hold;               // per member: 1 = not integrated this step
    unsigned char* h_hold = nullptr;        // pinned staging of d_hold
    unsigned int hold_cap = 0;
    bool any_hold = false;
    std::string last_error;
    ~esp_md_batch()
        {
        if (h_pred)
            (void)hipHostFree(h_pred);
        if (h_stage)
            (void)hipHostFree(h_stage);
        if (h_hold)
            (void)hipHostFree(h_hold);
        }
    };

namespace
{
int md_batch_fail(esp_md_batch* b, int code, const std::string& msg)
    {
    if (b)
        b->last_error = msg;
    g_md_error = msg;
    return code;
    }

template<class Fn> int md_batch_guard(esp_md_batch* b, Fn fn)
    {
    try
        {
        fn();
        return ESP_OK;
        }
    catch (const esp::EspError& e)
        {
        return md_batch_fail(b, e.code(), e.what());
        }
    catch (const std::exception& e)
        {
        return md_batch_fail(b, ESP_ERR_STATE, e.what());
        }
    }

//! Refresh system i's list pointers from its neighbour-list object.
void md_batch_refresh(esp_md_batch* b, unsigned int i)
    {
    esp_nlist* nl = b->nls[i];
    esp::md::MDSys& d = b->desc[i];
    if (nl && nl->built)
        {
        d.n_neigh = nl->n_neigh.as<unsigned int>();
        d.nlist = nl->nlist.as<unsigned int>();
        d.head = nl->head.as<uint64_t>();
        d.last_pos = nl->last_pos.as<double>();
        }
    else
        {
        d.n_neigh = nullptr;
        d.nlist = nullptr;
        d.head = nullptr;
        d.last_pos = nullptr;
        }
    if (nl && nl->built && nl->inner_built)
        {
        d.n_inner = nl->n_inner.as<unsigned int>();
        d.inner = nl->inner.as<unsigned int>();
        d.head_inner = nl->head_inner.as<uint64_t>();
        d.r_inner = nl->r_inner;
        }
    else
        {
        d.n_inner = nullptr;
        d.inner = nullptr;
        d.head_inner = nullptr;
        d.r_inner = 0.0;
        }
    }

void md_batch_upload(esp_md_batch* b)
    {
    esp::check_hip(hipMemcpy(b->d_desc.get(), b->desc.data(), b->desc.size() * sizeof(esp::md::MDSys), hipMemcpyHostToDevice),
                   "hipMemcpy descriptors");
    }
} // namespace

extern "C" {

int esp_md_batch_create(esp_md_batch** b)
    {
    if (!b)
        return md_batch_fail(nullptr, ESP_ERR_ARG, "esp_md_batch_create: null argument");
    *b = new esp_md_batch();
    return ESP_OK;
    }

int esp_md_batch_destroy(esp_md_batch* b)
    {
    delete b;
    return ESP_OK;
    }

const char* esp_md_batch_last_error(const esp_md_batch* b)
    {
    return b ? b->last_error.c_str() : g_md_error.c_str();
    }

static int md_batch_set_impl(esp_md_batch* b, unsigned int n, const esp_system* systems, double* const* vel, const esp_out* outs,
                             esp_nlist* const* nls, const uint64_t* seeds, bool async, hipStream_t stream);

int esp_md_batch_set(esp_md_batch* b, unsigned int n, const esp_system* systems, double* const* vel, const esp_out* outs,
                     esp_nlist* const* nls, const uint64_t* seeds)
    {
    return md_batch_set_impl(b, n, systems, vel, outs, nls, seeds, false, nullptr);
    }

int esp_md_batch_set_async(esp_md_batch* b, unsigned int n, const esp_system* systems, double* const* vel, const esp_out* outs,
                           esp_nlist* const* nls, const uint64_t* seeds, esp_stream_t stream)
    {
    return md_batch_set_impl(b, n, systems, vel, outs, nls, seeds, true, static_cast<hipStream_t>(stream));
    }

static int md_batch_set_impl(esp_md_batch* b, unsigned int n, const esp_system* systems, double* const* vel, const esp_out* outs,
                             esp_nlist* const* nls, const uint64_t* seeds, bool async, hipStream_t stream)
    {
    if (!b || (n && (!systems || !vel || !outs)))
        return md_batch_fail(b, ESP_ERR_ARG, "esp_md_batch_set: null argument");
    return md_batch_guard(b,
                          [&]
                          {
                              if (!b->pending.empty())
                                  throw esp::EspError(ESP_ERR_STATE, "esp_md_batch_set: asynchronous rebuilds are not finished");
                              b->any_hold = false;
                              b->sys.assign(systems, systems + n);
                              b->nls.assign(n, nullptr);
                              b->desc.assign(n, esp::md::MDSys {});
                              b->items.clear();
                              for (unsigned int i = 0; i < n; ++i)
                                  {
                                  const esp_system& s = systems[i];
                                  if (!s.postype || !vel[i] || !outs[i].force)
                                      throw esp::EspError(ESP_ERR_ARG, "esp_md_batch_set: positions, velocities and forces are required");
                                  if (s.box[3] != 0.0 || s.box[4] != 0.0 || s.box[5] != 0.0)
                                      throw esp::EspError(ESP_ERR_UNSUPPORTED, "triclinic boxes are not supported");
                                  if (nls)
                                      b->nls[i] = nls[i];
                                  esp::md::MDSys& d = b->desc[i];
                                  d.N = s.N;
                                  d.postype = const_cast<double*>(s.postype);
                                  d.vel = vel[i];
                                  d.force = outs[i].force;
                                  d.virial = outs[i].virial;
                                  d.pitch = outs[i].virial_pitch;
                                  for (int a = 0; a < 3; ++a)
                                      d.L[a] = s.box[a];
                                  d.seed = seeds ? seeds[i] : 0;
                                  md_batch_refresh(b, i);
                                  for (unsigned int c = 0; c < (s.N + 255u) / 256u; ++c)
                                      b->items.push_back({i, c});
                                  }
                              b->d_desc.resize(std::max<size_t>(n, 1) * sizeof(esp::md::MDSys));
                              b->d_items.resize(std::max<size_t>(b->items.size(), 1) * sizeof(esp::md::MDItem));
                              b->d_disp.resize(std::max<size_t>(n, 1) * 8);
                              b->h_disp.assign(n, 0ull);
                              const size_t dbytes = size_t(n) * sizeof(esp::md::MDSys);
                              const size_t ibytes = b->items.size() * sizeof(esp::md::MDItem);
                              if (!async)
                                  {
                                  if (n)
                                      md_batch_upload(b);
                                  if (ibytes)
                                      esp::check_hip(hipMemcpy(b->d_items.get(), b->items.data(), ibytes, hipMemcpyHostToDevice),
                                                     "hipMemcpy items");
                                  return;
                                  }
                              // through pinned staging on the caller's stream: the caller synchronises
                              // with that stream before the next set of this batch
                              if (b->stage_cap < dbytes + ibytes)
                                  {
                                  if (b->h_stage)
                                      esp::check_hip(hipHostFree(b->h_stage), "hipHostFree");
                                  b->h_stage = nullptr;
                                  b->stage_cap = 0;
                                  esp::check_hip(hipHostMalloc(&b->h_stage, (dbytes + ibytes) * 2, hipHostMallocDefault), "hipHostMalloc");
                                  b->stage_cap = (dbytes + ibytes) * 2;
                                  }
                              unsigned char* st = static_cast<unsigned char*>(b->h_stage);
                              if (dbytes)
                                  {
                                  std::memcpy(st, b->desc.data(), dbytes);
                                  esp::check_hip(hipMemcpyAsync(b->d_desc.get(), st, dbytes, hipMemcpyHostToDevice, stream),
                                                 "hipMemcpyAsync descriptors");
                                  }
                              if (ibytes)
                                  {
                                  std::memcpy(st + dbytes, b->items.data(), ibytes);
                                  esp::check_hip(hipMemcpyAsync(b->d_items.get(), st + dbytes, ibytes, hipMemcpyHostToDevice, stream),
                                                 "hipMemcpyAsync items");
                                  }
                          });
    }

int esp_md_batch_set_hold(esp_md_batch* b, const unsigned char* hold, esp_stream_t stream)
    {
    if (!b)
        return ESP_ERR_ARG;
    return md_batch_guard(b,
                          [&]
                          {
                              const unsigned int n = unsigned(b->desc.size());
                              bool any = false;
                              for (unsigned int i = 0; hold && i < n; ++i)
                                  any = any || hold[i] != 0;
                              b->any_hold = any;
                              if (!any)
                                  return;
                              if (b->hold_cap < n)
                                  {
                                  if (b->h_hold)
                                      esp::check_hip(hipHostFree(b->h_hold), "hipHostFree");
                                  b->h_hold = nullptr;
                                  b->hold_cap = 0;
                                  esp::check_hip(hipHostMalloc(reinterpret_cast<void**>(&b->h_hold), size_t(n) * 2, hipHostMallocDefault),
                                                 "hipHostMalloc");
                                  b->hold_cap = n * 2;
                                  }
                              b->d_hold.resize(std::max<size_t>(n, 1));
                              std::memcpy(b->h_hold, hold, n);
                              esp::check_hip(hipMemcpyAsync(b->d_hold.get(), b->h_hold, n, hipMemcpyHostToDevice, static_cast<hipStream_t>(stream)),
                                             "hipMemcpyAsync hold");
                          });
    }

int esp_md_batch_integrate_first(esp_md_batch* b, double dt, esp_stream_t stream)
    {
    if (!b)
        return ESP_ERR_ARG;
    return md_batch_guard(b,
                          [&]
                          {
                              esp::check_hip(esp::md::batch_vv_first(b->d_desc.as<esp::md::MDSys>(), b->d_items.as<esp::md::MDItem>(),
                                                                     unsigned(b->items.size()), dt, b->any_hold ? b->d_hold.as<unsigned char>() : nullptr,
                                                                     static_cast<hipStream_t>(stream)),
                                             "batch_vv_first");
                          });
    }

int esp_md_batch_integrate_second(esp_md_batch* b, double dt, double gamma, double kT, uint64_t step, esp_stream_t stream)
    {
    if (!b)
        return ESP_ERR_ARG;
    return md_batch_guard(b,
                          [&]
                          {
                              esp::check_hip(esp::md::batch_vv_second(b->d_desc.as<esp::md::MDSys>(), b->d_items.as<esp::md::MDItem>(),
                                                                      unsigned(b->items.size()), dt, gamma, kT, step,
                                                                      b->any_hold ? b->d_hold.as<unsigned char>() : nullptr,
                                                                      static_cast<hipStream_t>(stream)),
                                             "batch_vv_second");
                          });
    }

int esp_md_batch_wca(esp_md_batch* b, double sigma, double epsilon, esp_stream_t stream)
    {
    if (!b)
        return ESP_ERR_ARG;
    return md_batch_guard(b,
                          [&]
                          {
                              esp::check_hip(esp::md::batch_wca(b->d_desc.as<esp::md::MDSys>(), b->d_items.as<esp::md::MDItem>(),
                                                                unsigned(b->items.size()), sigma, epsilon, static_cast<hipStream_t>(stream)),
                                             "batch_wca");
                          });
    }

int esp_md_batch_predict(esp_md_batch* b, double dt, esp_stream_t stream)
    {
    if (!b)
        return ESP_ERR_ARG;
    return md_batch_guard(b,
                          [&]
                          {
                              hipStream_t s = static_cast<hipStream_t>(stream);
                              const unsigned int n = unsigned(b->desc.size());
                              if (n == 0)
                                  return;
                              if (b->pred_cap < n)
                                  {
                                  if (b->h_pred)
                                      esp::check_hip(hipHostFree(b->h_pred), "hipHostFree");
                                  b->h_pred = nullptr;
                                  esp::check_hip(hipHostMalloc(reinterpret_cast<void**>(&b->h_pred), size_t(n) * 8, hipHostMallocDefault),
                                                 "hipHostMalloc");
                                  b->pred_cap = n;
                                  }
                              b->d_disp_pred.resize(size_t(n) * 8);
                              esp::check_hip(hipMemsetAsync(b->d_disp_pred.get(), 0, size_t(n) * 8, s), "hipMemsetAsync");
                              esp::check_hip(esp::md::batch_predict_disp(b->d_desc.as<esp::md::MDSys>(), b->d_items.as<esp::md::MDItem>(),
                                                                         unsigned(b->items.size()), dt,
                                                                         b->d_disp_pred.as<unsigned long long>(), s),
                                             "batch_predict_disp");
                              esp::check_hip(hipMemcpyAsync(b->h_pred, b->d_disp_pred.get(), size_t(n) * 8, hipMemcpyDeviceToHost, s),
                                             "hipMemcpyAsync");
                          });
    }

//! Whether member i's list must be rebuilt once its positions have squared displacement d2 since
//! the last build (the criterion of esp_md_batch_update_nlists).
static bool needs_rebuild(const esp_md_batch* b, unsigned int i, double d2)
    {
    const esp_nlist* nl = b->nls[i];
    if (!nl)
        return false;
    const esp_system& sy = b->sys[i];
    return !nl->built || sy.N != nl->N || sy.box[0] != nl->box[0] || sy.box[1] != nl->box[1] || sy.box[2] != nl->box[2]
           || d2 >= 0.25 * nl->r_buff * nl->r_buff;
    }

int esp_md_batch_read_predicted(esp_md_batch* b, unsigned char* need)
    {
    if (!b || !need)
        return ESP_ERR_ARG;
    return md_batch_guard(b,
                          [&]
                          {
                              const unsigned int n = unsigned(b->desc.size());
                              if (n > b->pred_cap)
                                  throw esp::EspError(ESP_ERR_STATE, "esp_md_batch_read_predicted: no prediction for this member set");
                              for (unsigned int i = 0; i < n; ++i)
                                  {
                                  double d2;
                                  std::memcpy(&d2, &b->h_pred[i], 8);
                                  need[i] = needs_rebuild(b, i, d2) ? 1 : 0;
                                  }
                          });
    }

int esp_md_batch_rebuild_async(esp_md_batch* b, const unsigned char* which, unsigned char* moved, esp_stream_t stream)
    {
    if (!b || !which)
        return ESP_ERR_ARG;
    return md_batch_guard(b,
                          [&]
                          {
                              hipStream_t s = static_cast<hipStream_t>(stream);
                              const unsigned int n = unsigned(b->desc.size());
                              if (!b->pending.empty())
                                  throw esp::EspError(ESP_ERR_STATE, "esp_md_batch_rebuild_async: the previous rebuilds were not finished");
                              bool changed = false;
                              for (unsigned int i = 0; i < n; ++i)
                                  {
                                  if (moved)
                                      moved[i] = 0;
                                  esp_nlist* nl = b->nls[i];
                                  if (!which[i] || !nl)
                                      continue;
                                  const esp_system& sy = b->sys[i];
                                  const void* before[5] = {nl->n_neigh.get(), nl->nlist.get(), nl->head.get(), nl->inner.get(),
                                                           nl->last_pos.get()};
                                  build_prepare(nl, &sy);
                                  build_enqueue(nl, &sy, s);
                                  // positions at this build (the drift of the step has run): the
                                  // next predicted check measures from them
                                  esp::check_hip(esp::md::copy4(sy.N, sy.postype, nl->last_pos.as<double>(), s), "copy4");
                                  // taken as built; esp_md_batch_finish_rebuilds reports an overflow
                                  nl->inner_built = nl->want_inner;
                                  nl->built = true;
                                  nl->box[0] = sy.box[0];
                                  nl->box[1] = sy.box[1];
                                  nl->box[2] = sy.box[2];
                                  nl->builds += 1;
                                  b->pending.push_back(i);
                                  const void* after[5] = {nl->n_neigh.get(), nl->nlist.get(), nl->head.get(), nl->inner.get(),
                                                          nl->last_pos.get()};
                                  const bool mv = std::memcmp(before, after, sizeof(before)) != 0;
                                  if (moved)
                                      moved[i] = mv ? 1 : 0;
                                  esp::md::MDSys old = b->desc[i];
                                  md_batch_refresh(b, i);
                                  changed = changed || std::memcmp(&old, &b->desc[i], sizeof(old)) != 0;
                                  }
                              if (changed)
                                  esp::check_hip(hipMemcpyAsync(b->d_desc.get(), b->desc.data(), b->desc.size() * sizeof(esp::md::MDSys),
                                                                hipMemcpyHostToDevice, s),
                                                 "hipMemcpyAsync descriptors");
                          });
    }

int esp_md_batch_finish_rebuilds(esp_md_batch* b, unsigned char* overflowed, unsigned int* n_overflowed)
    {
    if (!b)
        return ESP_ERR_ARG;
    return md_batch_guard(b,
                          [&]
                          {
                              const unsigned int n = unsigned(b->desc.size());
                              unsigned int count = 0;
                              if (overflowed)
                                  std::memset(overflowed, 0, n);
                              for (unsigned int i : b->pending)
                                  {
                                  esp_nlist* nl = b->nls[i];
                                  const unsigned int over0 = nl->h_over[0], over1 = nl->h_over[1];
                                  if (over0 == 0 && over1 == 0)
                                      continue;
                                  // the rows were too short: raise the capacities, mark the list
                                  // unbuilt (the caller repeats the member's step with a fresh build)
                                  if (over0)
                                      nl->nmax = ((over0 + over0 / 4) + 7) & ~7u;
                                  if (over1)
                                      nl->nmax_in = ((over1 + over1 / 4) + 7) & ~7u;
                                  nl->built = false;
                                  nl->inner_built = false;
                                  if (overflowed && i < n)
                                      overflowed[i] = 1;
                                  ++count;
                                  }
                              b->pending.clear();
                              if (n_overflowed)
                                  *n_overflowed = count;
                          });
    }

int esp_md_batch_update_nlists(esp_md_batch* b, unsigned int* rebuilt, unsigned char* which, esp_stream_t stream)
    {
    if (!b)
        return ESP_ERR_ARG;
    return md_batch_guard(b,
                          [&]
                          {
                              hipStream_t s = static_cast<hipStream_t>(stream);
                              const unsigned int n = unsigned(b->desc.size());
                              unsigned int count = 0;
                              if (n)
                                  {
                                  esp::check_hip(hipMemsetAsync(b->d_disp.get(), 0, size_t(n) * 8, s), "hipMemsetAsync");
                                  esp::check_hip(esp::md::batch_max_disp(b->d_desc.as<esp::md::MDSys>(), b->d_items.as<esp::md::MDItem>(),
                                                                         unsigned(b->items.size()), b->d_disp.as<unsigned long long>(), s),
                                                 "batch_max_disp");
                                  esp::check_hip(hipMemcpyAsync(b->h_disp.data(), b->d_disp.get(), size_t(n) * 8, hipMemcpyDeviceToHost, s),
                                                 "hipMemcpyAsync");
                                  esp::check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
                                  }
                              // the lists to rebuild are enqueued together and waited for once; a
                              // list that overflowed its capacity is rebuilt on its own afterwards
                              std::vector<unsigned int> todo;
                              for (unsigned int i = 0; i < n; ++i)
                                  {
                                  esp_nlist* nl = b->nls[i];
                                  bool need = false;
                                  if (nl)
                                      {
                                      double d2;
                                      std::memcpy(&d2, &b->h_disp[i], 8);
                                      const esp_system& sy = b->sys[i];
                                      need = !nl->built || sy.N != nl->N || sy.box[0] != nl->box[0] || sy.box[1] != nl->box[1]
                                             || sy.box[2] != nl->box[2] || d2 >= 0.25 * nl->r_buff * nl->r_buff;
                                      if (need)
                                          {
                                          build_prepare(nl, &sy);
                                          build_enqueue(nl, &sy, s);
                                          todo.push_back(i);
                                          }
                                      }
                                  if (which)
                                      which[i] = need ? 1 : 0;
                                  }
                              bool changed = false;
                              if (!todo.empty())
                                  {
                                  esp::check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
                                  for (unsigned int i : todo)
                                      {
                                      esp_nlist* nl = b->nls[i];
                                      const esp_system& sy = b->sys[i];
                                      if (!build_finish(nl, &sy, s))
                                          do_build(nl, &sy, s);
                                      md_batch_refresh(b, i);
                                      changed = true;
                                      ++count;
                                      }
                                  }
                              if (changed)
                                  md_batch_upload(b);
                              if (rebuilt)
                                  *rebuilt = count;
                          });
    }


const char* esp_nlist_last_error(const esp_nlist* nl)
    {
    return nl ? nl->last_error.c_str() : g_md_error.c_str();
    }

int esp_nlist_build(esp_nlist* nl, const esp_system* sys, esp_stream_t stream)
    {
    if (!nl || !sys || !sys->postype)
        return fail(nl, ESP_ERR_ARG, "esp_nlist_build: null argument");
    return guard(nl, [&] { do_build(nl, sys, static_cast<hipStream_t>(stream)); });
    }

int esp_nlist_update(esp_nlist* nl, const esp_system* sys, int* rebuilt, esp_stream_t stream)
    {
    if (!nl || !sys || !sys->postype)
        return fail(nl, ESP_ERR_ARG, "esp_nlist_update: null argument");
    return guard(nl,
                 [&]
                 {
                     hipStream_t s = static_cast<hipStream_t>(stream);
                     bool need = !nl->built || sys->N != nl->N || sys->box[0] != nl->box[0] || sys->box[1] != nl->box[1]
                                 || sys->box[2] != nl->box[2];
                     if (!need)
                         {
                         esp::check_hip(hipMemsetAsync(nl->disp.get(), 0, 8, s), "hipMemsetAsync");
                         esp::check_hip(esp::md::max_displacement(sys->N, sys->postype, nl->last_pos.as<double>(), sys->box[0], sys->box[1],
                                                                  sys->box[2], nl->disp.as<unsigned long long>(), s),
                                        "max_displacement");
                         double d2 = 0.0;
                         esp::check_hip(hipMemcpyAsync(&d2, nl->disp.get(), 8, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
                         esp::check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
                         need = d2 >= 0.25 * nl->r_buff * nl->r_buff;
                         }
                     if (need)
                         do_build(nl, sys, s);
                     if (rebuilt)
                         *rebuilt = need ? 1 : 0;
                 });
    }

int esp_nlist_arrays(esp_nlist* nl, const unsigned int** n_neigh, const unsigned int** nlist, const uint64_t** head_list, unsigned int* nmax)
    {
    if (!nl || !nl->built)
        return fail(nl, ESP_ERR_STATE, "esp_nlist_arrays: list not built");
    if (n_neigh)
        *n_neigh = nl->n_neigh.as<unsigned int>();
    if (nlist)
        *nlist = nl->nlist.as<unsigned int>();
    if (head_list)
        *head_list = nl->head.as<uint64_t>();
    if (nmax)
        *nmax = nl->nmax;
    return ESP_OK;
    }

int esp_nlist_set_inner(esp_nlist* nl, double r_inner)
    {
    if (!nl || !(r_inner >= 0.0))
        return fail(nl, ESP_ERR_ARG, "esp_nlist_set_inner: r_inner must be >= 0");
    if (r_inner != nl->r_inner)
        {
        nl->r_inner = r_inner;
        nl->built = false;
        nl->inner_built = false;
        }
    return ESP_OK;
    }

int esp_nlist_set_cutoff(esp_nlist* nl, double r_cut)
    {
    if (!nl || !(r_cut > 0.0))
        return fail(nl, ESP_ERR_ARG, "esp_nlist_set_cutoff: r_cut must be > 0");
    if (r_cut != nl->r_cut)
        {
        nl->r_cut = r_cut;
        nl->built = false;
        nl->inner_built = false;
        }
    return ESP_OK;
    }

int esp_nlist_set_capacity(esp_nlist* nl, unsigned int nmax, unsigned int nmax_inner)
    {
    if (!nl)
        return fail(nl, ESP_ERR_ARG, "esp_nlist_set_capacity: null list");
    if (nmax == 0)
        {
        nl->fixed_capacity = false;
        return ESP_OK;
        }
    nl->fixed_capacity = true;
    nl->nmax = (nmax + 7u) & ~7u;
    nl->nmax_in = std::max(8u, (nmax_inner + 7u) & ~7u);
    return ESP_OK;
    }

int esp_nlist_inner_arrays(esp_nlist* nl, const unsigned int** n_neigh, const unsigned int** nlist, const uint64_t** head_list,
                           unsigned int* nmax)
    {
    if (!nl || !nl->built || !nl->inner_built)
        return fail(nl, ESP_ERR_STATE, "esp_nlist_inner_arrays: no inner list built (esp_nlist_set_inner with 0 < r_inner < r_cut)");
    if (n_neigh)
        *n_neigh = nl->n_inner.as<unsigned int>();
    if (nlist)
        *nlist = nl->inner.as<unsigned int>();
    if (head_list)
        *head_list = nl->head_inner.as<uint64_t>();
    if (nmax)
        *nmax = nl->nmax_in;
    return ESP_OK;
    }

int esp_nlist_stats(esp_nlist* nl, unsigned long long* builds)
    {
    if (!nl || !builds)
        return ESP_ERR_ARG;
    *builds = nl->builds;
    return ESP_OK;
    }

int esp_nlist_cell_order(esp_nlist* nl, const unsigned int** members, unsigned int* n)
    {
    if (!nl || !members || !n)
        return ESP_ERR_ARG;
    if (nl->builds == 0 || !nl->members.get())
        return fail(nl, ESP_ERR_STATE, "esp_nlist_cell_order: no list has been built");
    *members = nl->members.as<unsigned int>();
    *n = nl->N;
    return ESP_OK;
    }

int esp_gather_rows(unsigned int n, unsigned int row_bytes, const unsigned int* order, const void* src, void* dst,
                    esp_stream_t stream)
    {
    if (n == 0)
        return ESP_OK;
    if (!order || !src || !dst || src == dst || row_bytes == 0 || row_bytes % 4 != 0)
        return fail(nullptr, ESP_ERR_ARG, "esp_gather_rows: needs order, distinct src and dst, and row_bytes a multiple of 4");
    const hipError_t e = esp::md::gather_rows(n, row_bytes / 4, order, src, dst, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? ESP_OK : fail(nullptr, ESP_ERR_HIP, hipGetErrorString(e));
    }

int esp_wca_forces(const esp_system* sys, double sigma, double epsilon, esp_out* out, esp_stream_t stream)
    {
    if (!sys || !out || !out->force || !sys->nlist || !sys->n_neigh || !sys->head_list)
        return fail(nullptr, ESP_ERR_ARG, "esp_wca_forces: needs a system with a neighbour list and a force output");
    if (out->virial && out->virial_pitch < sys->N)
        return fail(nullptr, ESP_ERR_ARG, "esp_wca_forces: virial_pitch < N");
    const hipError_t e = esp::md::wca_forces(sys->N, sys->postype, sys->box[0], sys->box[1], sys->box[2], sys->n_neigh, sys->nlist,
                                             sys->head_list, sigma, epsilon, out->force, out->virial, out->virial_pitch,
                                             static_cast<hipStream_t>(stream));
    return e == hipSuccess ? ESP_OK : fail(nullptr, ESP_ERR_HIP, hipGetErrorString(e));
    }

int esp_integrate_first(unsigned int N, double* postype, double* vel, const double* force, double dt, const double* box, esp_stream_t stream)
    {
    if ((N > 0 && (!postype || !vel || !force)) || !box || !(dt > 0.0))
        return fail(nullptr, ESP_ERR_ARG, "esp_integrate_first: invalid argument");
    const hipError_t e = esp::md::vv_first(N, postype, vel, force, dt, box[0], box[1], box[2], static_cast<hipStream_t>(stream));
    return e == hipSuccess ? ESP_OK : fail(nullptr, ESP_ERR_HIP, hipGetErrorString(e));
    }

int esp_integrate_second(unsigned int N, double* vel, const double* force, double dt, double gamma, double kT, uint64_t seed,
                         uint64_t step, esp_stream_t stream)
    {
    if ((N > 0 && (!vel || !force)) || !(dt > 0.0) || gamma < 0.0 || kT < 0.0)
        return fail(nullptr, ESP_ERR_ARG, "esp_integrate_second: invalid argument");
    const hipError_t e = esp::md::vv_second(N, vel, force, dt, gamma, kT, seed, step, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? ESP_OK : fail(nullptr, ESP_ERR_HIP, hipGetErrorString(e));
    }

int esp_kinetic_potential(unsigned int N, const double* vel, const double* force, double* ke, double* pe)
    {
    if (!ke || !pe || (N > 0 && (!vel || !force)))
        return fail(nullptr, ESP_ERR_ARG, "esp_kinetic_potential: invalid argument");
    try
        {
        esp::DeviceBuffer part;
        const unsigned int nb = (N + 255) / 256;
        part.resize(size_t(std::max(nb, 1u)) * 16, true);
        esp::check_hip(esp::md::energies(N, vel, force, part.as<double>(), nullptr), "energies");
        std::vector<double> h(size_t(std::max(nb, 1u)) * 2, 0.0);
        esp::check_hip(hipMemcpy(h.data(), part.get(), h.size() * 8, hipMemcpyDeviceToHost), "hipMemcpy");
        double a = 0, b = 0;
        for (unsigned int i = 0; i < nb; ++i)
            {
            a += h[2 * i];
            b += h[2 * i + 1];
            }
        *ke = a;
        *pe = b;
        return ESP_OK;
        }
    catch (const esp::EspError& e)
        {
        return fail(nullptr, e.code(), e.what());
        }
    }

} // extern "C"
