// extern "C" entry points of include/esp_mi355x.h.  Every function catches all exceptions and
// converts them into a status code plus a per-context error message.
#include "../../include/esp_mi355x.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "esp_batch.h"
#include "esp_md.h"
#include "esp_solver.h"
#include "pswf_math.h"

struct esp_ctx
    {
    esp::EspSolver solver;
    std::string last_error;
    };

struct esp_batch
    {
    esp::EspBatch batch;
    std::string last_error;
    std::vector<esp::EspSolver*> solvers;
    std::vector<esp::SystemView> views;
    std::vector<esp::OutputView> outs;
    };

// ---- scratch of the no-allocation pair entry point ------------------------------------------
// One packed (x, y, z, q) record of 32 bytes per particle for up to ESP_NOALLOC_MAX_PARTICLES
// particles: 8 MiB that every GPU the library runs on holds until the process ends.  Host code
// never passes `scratch` itself to a kernel: its device address comes from hipGetSymbolAddress,
// once per device, with that device current.
static constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(ESP_NOALLOC_MAX_PARTICLES) * 32;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

static constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

//! Device address of `scratch` on the current device (looked up on the first call per device).
static void* scratch_on_current_device()
    {
    int dev = 0;
    esp::check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        throw esp::EspError(ESP_ERR_UNSUPPORTED, "device ordinal beyond the no-allocation scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev])
        {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw esp::EspError(ESP_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
        }
    return g_scratch_ptr[dev];
    }

namespace
{
thread_local std::string g_create_error;

esp::SolverParams to_solver(const esp_params* p)
    {
    esp::SolverParams s;
    s.nx = p->nx;
    s.ny = p->ny;
    s.nz = p->nz;
    s.order = p->order;
    s.kappa = p->kappa;
    s.rcut = p->rcut;
    s.alpha = p->alpha;
    s.n_table = p->n_table ? p->n_table : 4096u;
    s.precision = p->precision;
    s.split = p->split;
    s.window_c = p->window_c;
    s.window_degree = p->window_degree;
    s.pair_lanes = p->pair_lanes;
    s.k_cut = p->k_cut;
    s.tile = p->tile;
    return s;
    }

template<class Fn> int guarded_err(std::string* err, Fn fn)
    {
    std::string& msg = err ? *err : g_create_error;
    try
        {
        fn();
        if (err)
            err->clear();
        return ESP_OK;
        }
    catch (const esp::EspError& e)
        {
        msg = e.what();
        return e.code();
        }
    catch (const std::bad_alloc&)
        {
        msg = "host out of memory";
        return ESP_ERR_OOM;
        }
    catch (const std::invalid_argument& e)
        {
        msg = e.what();
        return ESP_ERR_ARG;
        }
    catch (const std::exception& e)
        {
        msg = e.what();
        return ESP_ERR_STATE;
        }
    catch (...)
        {
        msg = "unknown error";
        return ESP_ERR_STATE;
        }
    }

template<class Fn> int guarded(esp_ctx* ctx, Fn fn)
    {
    return guarded_err(ctx ? &ctx->last_error : nullptr, fn);
    }

esp::SystemView to_view(const esp_system* s)
    {
    esp::SystemView v;
    v.N = s->N;
    v.postype = s->postype;
    v.charge = s->charge;
    v.group_size = s->group_size;
    v.group_idx = s->group_idx;
    for (int i = 0; i < 6; ++i)
        v.box[i] = s->box[i];
    v.n_neigh = s->n_neigh;
    v.nlist = s->nlist;
    v.head_list = s->head_list;
    v.n_ex = s->n_ex;
    v.ex_list = s->ex_list;
    v.ex_pitch = s->ex_pitch;
    return v;
    }

esp::OutputView to_out(const esp_out* o)
    {
    esp::OutputView v;
    v.force = o->force;
    v.virial = o->virial;
    v.virial_pitch = o->virial_pitch;
    v.ext_energy_virial = o->ext_energy_virial;
    v.flags = o->flags;
    return v;
    }

esp::math::SplitKind kind_of(int split)
    {
    if (split == ESP_SPLIT_PSWF)
        return esp::math::SplitKind::PSWF;
    if (split == ESP_SPLIT_GAUSSIAN)
        return esp::math::SplitKind::GAUSSIAN;
    throw esp::EspError(ESP_ERR_ARG, "unknown split kind");
    }

int hip_status(hipError_t e)
    {
    if (e == hipSuccess)
        return ESP_OK;
    g_create_error = hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? ESP_ERR_OOM : ESP_ERR_HIP;
    }
} // namespace

extern "C" {

int esp_create(esp_ctx** ctx, const esp_params* params)
    {
    if (!ctx || !params)
        {
        g_create_error = "esp_create: null argument";
        return ESP_ERR_ARG;
        }
    *ctx = nullptr;
    esp_ctx* c = nullptr;
    const int rc = guarded(nullptr,
                           [&]
                           {
                               c = new esp_ctx();
                               try
                                   {
                                   c->solver.set_params(to_solver(params));
                                   }
                               catch (...)
                                   {
                                   delete c;
                                   c = nullptr;
                                   throw;
                                   }
                           });
    if (rc == ESP_OK)
        *ctx = c;
    return rc;
    }

int esp_set_params(esp_ctx* ctx, const esp_params* params)
    {
    if (!ctx || !params)
        return ESP_ERR_ARG;
    return guarded(ctx, [&] { ctx->solver.set_params(to_solver(params)); });
    }

int esp_invalidate(esp_ctx* ctx)
    {
    if (!ctx)
        return ESP_ERR_ARG;
    ctx->solver.invalidate();
    return ESP_OK;
    }

int esp_destroy(esp_ctx* ctx)
    {
    delete ctx;
    return ESP_OK;
    }

const char* esp_last_error(const esp_ctx* ctx)
    {
    return ctx ? ctx->last_error.c_str() : g_create_error.c_str();
    }

int esp_compute_forces(esp_ctx* ctx, uint64_t timestep, const esp_system* sys, esp_out* out, esp_stream_t stream)
    {
    (void)timestep;
    if (!ctx)
        return ESP_ERR_ARG;
    if (!sys || !out)
        {
        ctx->last_error = "esp_compute_forces: null system or output";
        return ESP_ERR_ARG;
        }
    return guarded(ctx, [&] { ctx->solver.compute(to_view(sys), to_out(out), static_cast<hipStream_t>(stream)); });
    }

int esp_prepare(esp_ctx* ctx, const esp_system* sys, const esp_out* out, esp_stream_t stream)
    {
    if (!ctx)
        return ESP_ERR_ARG;
    if (!sys || !out)
        {
        ctx->last_error = "esp_prepare: null system or output";
        return ESP_ERR_ARG;
        }
    return guarded(ctx, [&] { ctx->solver.setup(to_view(sys), to_out(out), static_cast<hipStream_t>(stream)); });
    }

int esp_pair_forces(esp_ctx* ctx, const esp_system* sys, esp_out* out, esp_stream_t stream)
    {
    if (!ctx)
        return ESP_ERR_ARG;
    if (!sys || !out)
        {
        ctx->last_error = "esp_pair_forces: null system or output";
        return ESP_ERR_ARG;
        }
    esp::OutputView o = to_out(out);
    o.flags = esp::COMPUTE_PAIR;
    return guarded(ctx, [&] { ctx->solver.compute(to_view(sys), o, static_cast<hipStream_t>(stream)); });
    }

int esp_pair_forces_noalloc(esp_ctx* ctx, const esp_system* sys, esp_out* out, esp_stream_t stream)
    {
    if (!ctx)
        return ESP_ERR_ARG;
    if (!sys || !out)
        {
        ctx->last_error = "esp_pair_forces_noalloc: null system or output";
        return ESP_ERR_ARG;
        }
    return guarded(ctx,
                   [&]
                   {
                       if (sys->N > ESP_NOALLOC_MAX_PARTICLES)
                           throw esp::EspError(ESP_ERR_RANGE, "esp_pair_forces_noalloc: N exceeds ESP_NOALLOC_MAX_PARTICLES");
                       void* p = scratch_on_current_device();
                       ctx->solver.pair_forces_into(to_view(sys), to_out(out), static_cast<hipStream_t>(stream), p,
                                                    NOALLOC_SCRATCH_BYTES);
                   });
    }

int esp_batch_create(esp_batch** b)
    {
    if (!b)
        return ESP_ERR_ARG;
    *b = nullptr;
    return guarded_err(nullptr, [&] { *b = new esp_batch(); });
    }

int esp_batch_destroy(esp_batch* b)
    {
    return guarded_err(nullptr, [&] { delete b; });
    }

const char* esp_batch_last_error(const esp_batch* b)
    {
    return b ? b->last_error.c_str() : g_create_error.c_str();
    }

int esp_batch_set_options(esp_batch* b, int use_graph, unsigned int max_grid)
    {
    if (!b)
        return ESP_ERR_ARG;
    return guarded_err(&b->last_error, [&] { b->batch.set_options(use_graph != 0, max_grid); });
    }

int esp_batch_compute(esp_batch* b, unsigned int n, esp_ctx* const* ctxs, const esp_system* systems, const esp_out* outs,
                      esp_stream_t stream)
    {
    if (!b)
        return ESP_ERR_ARG;
    if (n && (!ctxs || !systems || !outs))
        {
        b->last_error = "esp_batch_compute: null arrays";
        return ESP_ERR_ARG;
        }
    return guarded_err(&b->last_error,
                       [&]
                       {
                           b->solvers.resize(n);
                           b->views.resize(n);
                           b->outs.resize(n);
                           for (unsigned int i = 0; i < n; ++i)
                               {
                               if (!ctxs[i])
                                   throw esp::EspError(ESP_ERR_ARG, "esp_batch_compute: null context");
                               b->solvers[i] = &ctxs[i]->solver;
                               b->views[i] = to_view(&systems[i]);
                               b->outs[i] = to_out(&outs[i]);
                               }
                           b->batch.compute(n, b->solvers.data(), b->views.data(), b->outs.data(), static_cast<hipStream_t>(stream));
                       });
    }

int esp_batch_stats(esp_batch* b, unsigned long long* captures, unsigned long long* replays, unsigned long long* launches)
    {
    if (!b)
        return ESP_ERR_ARG;
    if (captures)
        *captures = b->batch.captures();
    if (replays)
        *replays = b->batch.replays();
    if (launches)
        *launches = b->batch.launches();
    return ESP_OK;
    }

int esp_get_info(esp_ctx* ctx, esp_info* info)
    {
    if (!ctx || !info)
        return ESP_ERR_ARG;
    return guarded(ctx,
                   [&]
                   {
                       const esp::SolverParams& p = ctx->solver.params();
                       std::memset(info, 0, sizeof(*info));
                       info->nx = p.nx;
                       info->ny = p.ny;
                       info->nz = p.nz;
                       info->order = p.order;
                       info->kappa = p.kappa;
                       info->rcut = p.rcut;
                       info->alpha = p.alpha;
                       info->n_table = p.n_table;
                       info->precision = p.precision;
                       info->split = p.split;
                       ctx->solver.build_host_tables();
                       info->window_degree = ctx->solver.window().D;
                       info->window_c = ctx->solver.window().cw;
                       info->self_coeff = ctx->solver.self_coeff();
                       info->background_coeff = ctx->solver.background_coeff();
                       info->tile = unsigned(ctx->solver.geometry().T);
                       info->device_bytes = ctx->solver.memory_bytes();
                       info->k_cut = ctx->solver.effective_k_cut();
                   });
    }

int esp_get_charge_sums(esp_ctx* ctx, double* q_sum, double* q2_sum)
    {
    if (!ctx || !q_sum || !q2_sum)
        return ESP_ERR_ARG;
    return guarded(ctx, [&] { ctx->solver.charge_sums(*q_sum, *q2_sum); });
    }

int esp_get_table_size(esp_ctx* ctx, unsigned int* n)
    {
    if (!ctx || !n)
        return ESP_ERR_ARG;
    *n = ctx->solver.params().n_table;
    return ESP_OK;
    }

int esp_get_table_entry(esp_ctx* ctx, unsigned int i, esp_table_entry* entry)
    {
    if (!ctx || !entry)
        return ESP_ERR_ARG;
    return guarded(ctx,
                   [&]
                   {
                       const esp::SolverParams& p = ctx->solver.params();
                       if (i >= p.n_table)
                           throw esp::EspError(ESP_ERR_RANGE, "ESPForceCompute table index out of range");
                       // buildPSWFTable (ESPForceCompute.cc:336-372): L and -dL/dr at segment midpoints,
                       // constant-segment entries
                       esp::math::SplitFunction split(p.split == 1 ? esp::math::SplitKind::GAUSSIAN : esp::math::SplitKind::PSWF,
                                                      p.kappa, p.rcut);
                       const double inv_seg = 1.0 / double(p.n_table);
                       const double r_lo = double(i) * inv_seg * p.rcut;
                       const double r_hi = double(i + 1) * inv_seg * p.rcut;
                       const double r_mid = std::max(0.5 * (r_lo + r_hi), 1.0e-9);
                       const double s = r_mid / p.rcut;
                       const double S = split.S(s);
                       const double dS = split.dSds(s) / p.rcut;
                       const double inv_4pi_r = esp::math::INV_4PI / r_mid;
                       std::memset(entry, 0, sizeof(*entry));
                       entry->potential_coeffs[0] = S * inv_4pi_r;
                       entry->force_coeffs[0] = S * inv_4pi_r / r_mid - dS * inv_4pi_r;
                       entry->r_lo = r_lo;
                       entry->dr_inv = (r_hi > r_lo) ? 1.0 / (r_hi - r_lo) : 0.0;
                   });
    }

int esp_get_split_table(esp_ctx* ctx, unsigned int* n_segs, double* screen, double* dscreen, double* r_lo, double* dr_inv)
    {
    if (!ctx || !n_segs)
        return ESP_ERR_ARG;
    return guarded(ctx,
                   [&]
                   {
                       if (ctx->solver.split_r_lo().empty())
                           ctx->solver.build_host_tables();
                       const unsigned int n = unsigned(ctx->solver.split_r_lo().size());
                       *n_segs = n;
                       if (screen)
                           std::memcpy(screen, ctx->solver.split_screen().data(), sizeof(double) * 5 * n);
                       if (dscreen)
                           std::memcpy(dscreen, ctx->solver.split_dscreen().data(), sizeof(double) * 5 * n);
                       if (r_lo)
                           std::memcpy(r_lo, ctx->solver.split_r_lo().data(), sizeof(double) * n);
                       if (dr_inv)
                           std::memcpy(dr_inv, ctx->solver.split_dr_inv().data(), sizeof(double) * n);
                   });
    }

int esp_get_fixed_point_scale(esp_ctx* ctx, float* scale, int* lowered)
    {
    if (!ctx || !scale || !lowered)
        return ESP_ERR_ARG;
    return guarded(ctx, [&] { ctx->solver.fixed_point_scale(*scale, *lowered); });
    }

int esp_get_pair_table(esp_ctx* ctx, const double** host_table, const double** device_table, unsigned int* n_segs)
    {
    if (!ctx || !host_table || !device_table || !n_segs)
        return ESP_ERR_ARG;
    return guarded(ctx, [&] { ctx->solver.pair_table(host_table, device_table, n_segs); });
    }

int esp_set_split_table(esp_ctx* ctx, unsigned int n_segs, const double* screen, const double* dscreen, const double* r_lo,
                        const double* dr_inv)
    {
    if (!ctx)
        return ESP_ERR_ARG;
    return guarded(ctx, [&] { ctx->solver.set_split_table(n_segs, screen, dscreen, r_lo, dr_inv); });
    }

int esp_get_window(esp_ctx* ctx, unsigned int* P, unsigned int* D, double* coef)
    {
    if (!ctx || !P || !D)
        return ESP_ERR_ARG;
    return guarded(ctx,
                   [&]
                   {
                       if (ctx->solver.window().coef.empty())
                           ctx->solver.build_host_tables();
                       const esp::math::WindowTable& w = ctx->solver.window();
                       *P = w.P;
                       *D = w.D;
                       if (coef)
                           std::memcpy(coef, w.coef.data(), sizeof(double) * w.coef.size());
                   });
    }

// ---- tracing ------------------------------------------------------------------------------
int esp_set_graph(esp_ctx* ctx, int enable)
    {
    if (!ctx || enable < 0 || enable > 1)
        return ESP_ERR_ARG;
    return guarded(ctx, [&] { ctx->solver.set_graph(enable != 0); });
    }

int esp_get_graph_stats(esp_ctx* ctx, unsigned long long* captures, unsigned long long* replays)
    {
    if (!ctx || !captures || !replays)
        return ESP_ERR_ARG;
    ctx->solver.graph_stats(*captures, *replays);
    return ESP_OK;
    }

int esp_set_profiling(esp_ctx* ctx, int enable)
    {
    if (!ctx)
        return ESP_ERR_ARG;
    if (enable < 0 || enable > 2)
        return ESP_ERR_ARG;
    ctx->solver.set_profiling(enable);
    return ESP_OK;
    }

int esp_get_stage_times(esp_ctx* ctx, double* ms, unsigned int* calls)
    {
    if (!ctx || !ms || !calls)
        return ESP_ERR_ARG;
    return guarded(ctx, [&] { ctx->solver.stage_times(ms, calls); });
    }

int esp_reset_stage_times(esp_ctx* ctx)
    {
    if (!ctx)
        return ESP_ERR_ARG;
    return guarded(ctx, [&] { ctx->solver.reset_stage_times(); });
    }

const char* esp_stage_name(int i)
    {
    return esp::EspSolver::stage_name(i);
    }

// ---- host-only mathematics -------------------------------------------------------------
int esp_build_split_table(int split, double kappa, double rcut, unsigned int n_segs, double* screen, double* dscreen, double* r_lo,
                          double* dr_inv)
    {
    if (!screen || !dscreen || !r_lo || !dr_inv || n_segs < 1)
        return ESP_ERR_ARG;
    return guarded(nullptr,
                   [&]
                   {
                       esp::math::SplitFunction f(kind_of(split), kappa, rcut);
                       std::vector<double> a, b, c, d;
                       esp::math::build_split_table(f, n_segs, a, b, c, d);
                       std::memcpy(screen, a.data(), sizeof(double) * a.size());
                       std::memcpy(dscreen, b.data(), sizeof(double) * b.size());
                       std::memcpy(r_lo, c.data(), sizeof(double) * c.size());
                       std::memcpy(dr_inv, d.data(), sizeof(double) * d.size());
                   });
    }

int esp_split_eval(int split, double kappa, double rcut, int n, const double* s, double* S, double* dSds)
    {
    if (n < 0 || (n > 0 && (!s || !S || !dSds)))
        return ESP_ERR_ARG;
    return guarded(nullptr,
                   [&]
                   {
                       esp::math::SplitFunction f(kind_of(split), kappa, rcut);
                       for (int i = 0; i < n; ++i)
                           {
                           S[i] = f.S(s[i]);
                           dSds[i] = f.dSds(s[i]);
                           }
                   });
    }

int esp_split_fourier(int split, double kappa, double rcut, int n, const double* k, double* F, double* dF)
    {
    if (n < 0 || (n > 0 && (!k || !F || !dF)))
        return ESP_ERR_ARG;
    return guarded(nullptr,
                   [&]
                   {
                       esp::math::SplitFunction f(kind_of(split), kappa, rcut);
                       for (int i = 0; i < n; ++i)
                           {
                           F[i] = f.F(k[i]);
                           dF[i] = f.dF(k[i]);
                           }
                   });
    }

int esp_split_constants(int split, double kappa, double rcut, double* self_coeff, double* background_coeff)
    {
    if (!self_coeff || !background_coeff)
        return ESP_ERR_ARG;
    return guarded(nullptr,
                   [&]
                   {
                       esp::math::SplitFunction f(kind_of(split), kappa, rcut);
                       *self_coeff = f.self_coeff();
                       *background_coeff = f.background_coeff();
                   });
    }

int esp_prolate_eval(double c, int n, const double* x, double* psi, double* integral)
    {
    if (n < 0 || (n > 0 && (!x || !psi || !integral)))
        return ESP_ERR_ARG;
    return guarded(nullptr,
                   [&]
                   {
                       esp::math::Prolate0 p(c);
                       for (int i = 0; i < n; ++i)
                           {
                           psi[i] = p.eval(x[i]);
                           integral[i] = p.integral(x[i]);
                           }
                   });
    }

int esp_build_window(unsigned int P, unsigned int D, double window_c, double* coef)
    {
    if (!coef)
        return ESP_ERR_ARG;
    return guarded(nullptr,
                   [&]
                   {
                       esp::math::WindowTable w = esp::math::build_window(P, D, window_c);
                       std::memcpy(coef, w.coef.data(), sizeof(double) * w.coef.size());
                   });
    }

int esp_window_fourier(unsigned int P, unsigned int D, const double* coef, int n, const double* theta, double* what)
    {
    if (!coef || n < 0 || (n > 0 && (!theta || !what)) || P < 2 || P > 16 || D > 15)
        return ESP_ERR_ARG;
    esp::math::WindowTable w;
    w.P = P;
    w.D = D;
    w.coef.assign(coef, coef + size_t(P) * (D + 1));
    for (int i = 0; i < n; ++i)
        what[i] = w.fourier(theta[i]);
    return ESP_OK;
    }

double esp_default_window_c(unsigned int P, double sigma)
    {
    return esp::math::default_window_c(P, sigma);
    }

unsigned int esp_default_window_degree(unsigned int P)
    {
    return esp::math::default_window_degree(P);
    }

// ---- device utilities ---------------------------------------------------------------------
int esp_device_count(int* n)
    {
    if (!n)
        return ESP_ERR_ARG;
    int c = 0;
    const hipError_t e = hipGetDeviceCount(&c);
    *n = (e == hipSuccess) ? c : 0;
    return ESP_OK;
    }

int esp_set_device(int device)
    {
    return hip_status(hipSetDevice(device));
    }

int esp_device_malloc(void** ptr, size_t bytes)
    {
    if (!ptr)
        return ESP_ERR_ARG;
    return hip_status(hipMalloc(ptr, bytes));
    }

int esp_device_free(void* ptr)
    {
    return hip_status(hipFree(ptr));
    }

int esp_memcpy_htod(void* dst, const void* src, size_t bytes)
    {
    return hip_status(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    }

int esp_memcpy_dtoh(void* dst, const void* src, size_t bytes)
    {
    return hip_status(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    }

int esp_memcpy_htod_async(void* dst, const void* src, size_t bytes, esp_stream_t stream)
    {
    return hip_status(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, static_cast<hipStream_t>(stream)));
    }

int esp_host_malloc_pinned(void** ptr, size_t bytes)
    {
    if (!ptr)
        return ESP_ERR_ARG;
    *ptr = nullptr;
    return hip_status(hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault));
    }

int esp_host_free_pinned(void* ptr)
    {
    return ptr ? hip_status(hipHostFree(ptr)) : ESP_OK;
    }

int esp_memcpy_dtod(void* dst, const void* src, size_t bytes, esp_stream_t stream)
    {
    return hip_status(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, static_cast<hipStream_t>(stream)));
    }

int esp_memset(void* dst, int value, size_t bytes)
    {
    return hip_status(hipMemset(dst, value, bytes));
    }

int esp_copy_records(void* dst, const void* src, unsigned int n_records, esp_stream_t stream)
    {
    if ((!dst || !src) && n_records)
        return ESP_ERR_ARG;
    return hip_status(esp::md::copy4(n_records, static_cast<const double*>(src), static_cast<double*>(dst),
                                     static_cast<hipStream_t>(stream)));
    }

int esp_stream_create(esp_stream_t* stream)
    {
    if (!stream)
        return ESP_ERR_ARG;
    hipStream_t s = nullptr;
    const int rc = hip_status(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return rc;
    }

int esp_stream_destroy(esp_stream_t stream)
    {
    return hip_status(hipStreamDestroy(static_cast<hipStream_t>(stream)));
    }

int esp_stream_synchronize(esp_stream_t stream)
    {
    return hip_status(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    }

int esp_device_synchronize(void)
    {
    return hip_status(hipDeviceSynchronize());
    }

int esp_event_create(void** ev)
    {
    if (!ev)
        return ESP_ERR_ARG;
    hipEvent_t e = nullptr;
    const int rc = hip_status(hipEventCreate(&e));
    *ev = e;
    return rc;
    }

int esp_event_destroy(void* ev)
    {
    return hip_status(hipEventDestroy(static_cast<hipEvent_t>(ev)));
    }

int esp_event_record(void* ev, esp_stream_t stream)
    {
    return hip_status(hipEventRecord(static_cast<hipEvent_t>(ev), static_cast<hipStream_t>(stream)));
    }

int esp_event_elapsed_ms(void* start, void* stop, float* ms)
    {
    if (!ms)
        return ESP_ERR_ARG;
    return hip_status(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
    }

const char* esp_version(void)
    {
    return "esp_mi355x 0.1.0";
    }

} // extern "C"
