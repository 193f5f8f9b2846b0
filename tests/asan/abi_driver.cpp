// Host AddressSanitizer driver for the C-ABI (built by esp_hoomd_amd.build.build_asan with the
// library's host sources instrumented: -Xarch_host -fsanitize=address; device code is not).
//
// It walks every entry point family through the library's own header -- host table math,
// context lifetime and introspection, argument errors, and on a GPU the force, pair, batch,
// neighbour-list and MD paths -- so heap overflows, use-after-free and leaks in the host code
// (the argument handling, table builders, buffer management and error paths) are reported by
// ASan / LeakSanitizer.  The checks it makes on results are invariants (finite values, Newton's
// third law, bit-identical FP64 repeats and batch-vs-alone results, no-allocation pair path
// equal to the allocating one); numerical parity lives in the pytest suites.
//
// Usage: abi_driver [--host-only]   (exit status 0 = every check passed)
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include <sanitizer/lsan_interface.h>
#include <unistd.h>

#include "esp_mi355x.h"

static int g_fail = 0;
#define CHECK(cond)                                                                        \
    do                                                                                     \
        {                                                                                  \
        if (!(cond))                                                                       \
            {                                                                              \
            std::fprintf(stderr, "CHECK failed at line %d: %s\n", __LINE__, #cond);        \
            ++g_fail;                                                                      \
            }                                                                              \
        } while (0)
#define OK(call) CHECK((call) == ESP_OK)

static bool all_finite(const std::vector<double>& v)
    {
    for (double x : v)
        if (!std::isfinite(x))
            return false;
    return true;
    }

static esp_params make_params(unsigned int n, unsigned int P, double kappa, double rcut, int precision)
    {
    esp_params p;
    std::memset(&p, 0, sizeof(p));
    p.nx = p.ny = p.nz = n;
    p.order = P;
    p.kappa = kappa;
    p.rcut = rcut;
    p.n_table = 4096;
    p.precision = precision;
    p.split = ESP_SPLIT_PSWF;
    return p;
    }

// ---- host-only entry points -------------------------------------------------------------
static void host_tables()
    {
    for (int split : {ESP_SPLIT_PSWF, ESP_SPLIT_GAUSSIAN})
        for (unsigned int n : {16u, 512u, 4096u})
            {
            std::vector<double> S(5 * n), dS(5 * n), lo(n), dri(n);
            OK(esp_build_split_table(split, 3.0, 4.0, n, S.data(), dS.data(), lo.data(), dri.data()));
            CHECK(all_finite(S) && all_finite(dS) && all_finite(lo) && all_finite(dri));
            }
    std::vector<double> s(257), S(257), dS(257);
    for (int i = 0; i < 257; ++i)
        s[i] = i / 256.0;
    OK(esp_split_eval(ESP_SPLIT_PSWF, 3.0, 4.0, 257, s.data(), S.data(), dS.data()));
    CHECK(std::fabs(S[0] - 1.0) < 1e-12 && std::fabs(S[256]) < 1e-6);
    std::vector<double> k(100), F(100), dF(100);
    for (int i = 0; i < 100; ++i)
        k[i] = 0.05 * i;
    OK(esp_split_fourier(ESP_SPLIT_PSWF, 3.0, 4.0, 100, k.data(), F.data(), dF.data()));
    CHECK(all_finite(F) && all_finite(dF));
    double self = 0, bg = 0;
    OK(esp_split_constants(ESP_SPLIT_PSWF, 3.0, 4.0, &self, &bg));
    CHECK(self > 0 && std::isfinite(bg));
    std::vector<double> x(101), psi(101), integral(101);
    for (int i = 0; i < 101; ++i)
        x[i] = -1.0 + i / 50.0;
    OK(esp_prolate_eval(12.0, 101, x.data(), psi.data(), integral.data()));
    CHECK(all_finite(psi) && integral[100] > 0 && std::fabs(integral[0] + integral[100]) < 1e-12);
    for (unsigned int P = 2; P <= 16; ++P)
        {
        const unsigned int D = esp_default_window_degree(P);
        std::vector<double> coef(size_t(P) * (D + 1));
        OK(esp_build_window(P, D, esp_default_window_c(P, 1.25), coef.data()));
        CHECK(all_finite(coef));
        std::vector<double> th(64), what(64);
        for (int i = 0; i < 64; ++i)
            th[i] = -3.14159 + i * 0.1;
        OK(esp_window_fourier(P, D, coef.data(), 64, th.data(), what.data()));
        CHECK(all_finite(what));
        }
    // argument errors come back as codes, never as crashes
    CHECK(esp_build_split_table(ESP_SPLIT_PSWF, 3.0, 4.0, 0, nullptr, nullptr, nullptr, nullptr) != ESP_OK);
    CHECK(esp_build_window(0, 4, 1.0, nullptr) != ESP_OK);
    CHECK(esp_build_window(17, 4, 1.0, nullptr) != ESP_OK);
    CHECK(esp_split_eval(ESP_SPLIT_PSWF, -1.0, 4.0, 257, s.data(), S.data(), dS.data()) != ESP_OK);
    CHECK(esp_prolate_eval(12.0, 4, nullptr, nullptr, nullptr) != ESP_OK);
    }

static void context_host_paths()
    {
    esp_ctx* ctx = nullptr;
    CHECK(esp_create(nullptr, nullptr) == ESP_ERR_ARG);
    CHECK(esp_create(&ctx, nullptr) == ESP_ERR_ARG && ctx == nullptr);
    esp_params bad = make_params(64, 1, 2.0, 3.0, ESP_PREC_FP64);
    CHECK(esp_create(&ctx, &bad) != ESP_OK && ctx == nullptr);
    bad = make_params(64, 8, 2.0, 3.0, ESP_PREC_FP64);
    bad.alpha = -0.5;
    CHECK(esp_create(&ctx, &bad) == ESP_ERR_ARG);
    CHECK(std::strlen(esp_last_error(nullptr)) > 0);

    esp_params p = make_params(64, 8, 3.0, 4.0, ESP_PREC_FP64);
    OK(esp_create(&ctx, &p));
    CHECK(ctx != nullptr);
    if (!ctx)
        return;
    unsigned int n = 0;
    OK(esp_get_split_table(ctx, &n, nullptr, nullptr, nullptr, nullptr));
    CHECK(n == 4096);
    std::vector<double> S(5 * size_t(n)), dS(5 * size_t(n)), lo(n), dri(n);
    OK(esp_get_split_table(ctx, &n, S.data(), dS.data(), lo.data(), dri.data()));
    OK(esp_set_split_table(ctx, n, S.data(), dS.data(), lo.data(), dri.data()));
    CHECK(esp_set_split_table(ctx, 0, S.data(), dS.data(), lo.data(), dri.data()) != ESP_OK);
    unsigned int nt = 0;
    OK(esp_get_table_size(ctx, &nt));
    for (unsigned int i = 0; i < nt; ++i)
        {
        esp_table_entry e;
        OK(esp_get_table_entry(ctx, i, &e));
        }
    esp_table_entry e;
    CHECK(esp_get_table_entry(ctx, nt, &e) != ESP_OK);
    CHECK(std::strlen(esp_last_error(ctx)) > 0);
    unsigned int P = 0, D = 0;
    OK(esp_get_window(ctx, &P, &D, nullptr));
    std::vector<double> coef(size_t(P) * (D + 1));
    OK(esp_get_window(ctx, &P, &D, coef.data()));
    CHECK(P == 8 && all_finite(coef));
    esp_info info;
    OK(esp_get_info(ctx, &info));
    CHECK(info.order == 8 && info.nx == 64);
    esp_params p2 = make_params(48, 6, 2.5, 4.0, ESP_PREC_MIXED);
    OK(esp_set_params(ctx, &p2));
    p2.nx = 1;
    CHECK(esp_set_params(ctx, &p2) != ESP_OK);
    OK(esp_invalidate(ctx));
    for (int i = 0; i < ESP_N_STAGES; ++i)
        CHECK(esp_stage_name(i) != nullptr);
    OK(esp_destroy(ctx));
    }

// ---- GPU paths ---------------------------------------------------------------------------
struct Dev
    {
    void* p = nullptr;
    explicit Dev(size_t bytes) { OK(esp_device_malloc(&p, bytes)); }
    ~Dev() { (void)esp_device_free(p); }
    template<typename T> T* as() const { return static_cast<T*>(p); }
    };

static void gpu_paths()
    {
    const unsigned int N = 4096;
    const double L = 16.0;
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> u(-L / 2, L / 2);
    std::vector<double> pos(4 * N), q(N), vel(4 * N, 0.0);
    for (unsigned int i = 0; i < N; ++i)
        {
        for (int a = 0; a < 3; ++a)
            pos[4 * i + a] = u(rng);
        pos[4 * i + 3] = 0.0;
        q[i] = (i % 2) ? -1.0 : 1.0;
        vel[4 * i + 3] = 1.0;
        }
    Dev d_pos(pos.size() * 8), d_q(N * 8), d_vel(vel.size() * 8), d_f(4 * N * 8), d_f2(4 * N * 8), d_vir(6 * N * 8), d_ext(7 * 8);
    OK(esp_memcpy_htod(d_pos.p, pos.data(), pos.size() * 8));
    OK(esp_memcpy_htod(d_q.p, q.data(), N * 8));
    OK(esp_memcpy_htod(d_vel.p, vel.data(), vel.size() * 8));

    esp_system sys;
    std::memset(&sys, 0, sizeof(sys));
    sys.N = N;
    sys.postype = d_pos.as<double>();
    sys.charge = d_q.as<double>();
    sys.box[0] = sys.box[1] = sys.box[2] = L;

    esp_nlist* nl = nullptr;
    OK(esp_nlist_create(&nl, 3.5, 0.3));
    OK(esp_nlist_set_inner(nl, 1.12246));
    OK(esp_nlist_build(nl, &sys, nullptr));
    unsigned int nmax = 0;
    OK(esp_nlist_arrays(nl, &sys.n_neigh, &sys.nlist, &sys.head_list, &nmax));
    int rebuilt = -1;
    OK(esp_nlist_update(nl, &sys, &rebuilt, nullptr));
    CHECK(rebuilt == 0);

    esp_out out;
    std::memset(&out, 0, sizeof(out));
    out.force = d_f.as<double>();
    out.virial = d_vir.as<double>();
    out.virial_pitch = N;
    out.ext_energy_virial = d_ext.as<double>();

    // FP64: two evaluations are bit-identical; the forces sum to ~0
    esp_params p = make_params(40, 8, 3.0, 3.5, ESP_PREC_FP64);
    esp_ctx* ctx = nullptr;
    OK(esp_create(&ctx, &p));
    OK(esp_set_profiling(ctx, 1));
    OK(esp_compute_forces(ctx, 0, &sys, &out, nullptr));
    std::vector<double> f1(4 * N), f2(4 * N);
    OK(esp_memcpy_dtoh(f1.data(), d_f.p, f1.size() * 8));
    OK(esp_compute_forces(ctx, 1, &sys, &out, nullptr));
    OK(esp_memcpy_dtoh(f2.data(), d_f.p, f2.size() * 8));
    CHECK(std::memcmp(f1.data(), f2.data(), f1.size() * 8) == 0);
    CHECK(all_finite(f1));
    {
    // set-up ahead of the first computation (stream-ordered uploads from pinned staging, on
    // another stream than the computation) changes nothing; so does a second set-up after new
    // parameters, which reuses the staging
    esp_ctx* pre = nullptr;
    OK(esp_create(&pre, &p));
    esp_stream_t st = nullptr;
    OK(esp_stream_create(&st));
    OK(esp_prepare(pre, &sys, &out, st));
    OK(esp_compute_forces(pre, 0, &sys, &out, nullptr));
    std::vector<double> f3(4 * N);
    OK(esp_stream_synchronize(nullptr));
    OK(esp_memcpy_dtoh(f3.data(), d_f.p, f3.size() * 8));
    CHECK(std::memcmp(f1.data(), f3.data(), f1.size() * 8) == 0);
    esp_params p_other = make_params(32, 6, 3.0, 3.5, ESP_PREC_FP64);
    OK(esp_set_params(pre, &p_other));
    OK(esp_prepare(pre, &sys, &out, st));
    OK(esp_set_params(pre, &p));
    OK(esp_prepare(pre, &sys, &out, st));
    OK(esp_compute_forces(pre, 0, &sys, &out, st));
    OK(esp_stream_synchronize(st));
    OK(esp_memcpy_dtoh(f3.data(), d_f.p, f3.size() * 8));
    CHECK(std::memcmp(f1.data(), f3.data(), f1.size() * 8) == 0);
    CHECK(esp_prepare(pre, nullptr, &out, st) == ESP_ERR_ARG);
    OK(esp_destroy(pre));
    OK(esp_stream_destroy(st));
    }
    {
    // cell order of the list's last build is a permutation; gathering rows by it and back
    // (by its inverse) restores the positions exactly
    const unsigned int* order = nullptr;
    unsigned int n_order = 0;
    OK(esp_nlist_cell_order(nl, &order, &n_order));
    CHECK(n_order == N);
    std::vector<unsigned int> h_order(N);
    OK(esp_memcpy_dtoh(h_order.data(), order, N * 4));
    std::vector<unsigned int> seen(N, 0), inverse(N, 0);
    for (unsigned int i = 0; i < N; ++i)
        if (h_order[i] < N)
            {
            ++seen[h_order[i]];
            inverse[h_order[i]] = i;
            }
    bool perm = true;
    for (unsigned int i = 0; i < N; ++i)
        perm = perm && seen[i] == 1;
    CHECK(perm);
    Dev d_sorted(pos.size() * 8), d_back(pos.size() * 8), d_inv(N * 4);
    OK(esp_memcpy_htod(d_inv.p, inverse.data(), N * 4));
    OK(esp_gather_rows(N, 32, order, d_pos.p, d_sorted.p, nullptr));
    OK(esp_gather_rows(N, 32, d_inv.as<unsigned int>(), d_sorted.p, d_back.p, nullptr));
    std::vector<double> back(pos.size());
    OK(esp_memcpy_dtoh(back.data(), d_back.p, back.size() * 8));
    CHECK(std::memcmp(back.data(), pos.data(), pos.size() * 8) == 0);
    CHECK(esp_gather_rows(N, 6, order, d_pos.p, d_sorted.p, nullptr) == ESP_ERR_ARG);
    CHECK(esp_gather_rows(N, 32, order, d_pos.p, d_pos.p, nullptr) == ESP_ERR_ARG);
    }
    double fs[3] = {0, 0, 0}, fa = 0;
    for (unsigned int i = 0; i < N; ++i)
        for (int a = 0; a < 3; ++a)
            {
            fs[a] += f1[4 * i + a];
            fa += std::fabs(f1[4 * i + a]);
            }
    CHECK(std::fabs(fs[0]) + std::fabs(fs[1]) + std::fabs(fs[2]) < 1e-8 * fa);
    double ms[ESP_N_STAGES];
    unsigned int calls[ESP_N_STAGES];
    OK(esp_get_stage_times(ctx, ms, calls));
    unsigned int most = 0;
    for (int i = 0; i < ESP_N_STAGES; ++i)
        most = calls[i] > most ? calls[i] : most;
    CHECK(most == 2);
    OK(esp_reset_stage_times(ctx));
    double qs = 0, q2 = 0;
    OK(esp_get_charge_sums(ctx, &qs, &q2));
    CHECK(std::fabs(qs) < 1e-12 && std::fabs(q2 - N) < 1e-9);
    float fx = 0;
    int lowered = 1;
    OK(esp_get_fixed_point_scale(ctx, &fx, &lowered));
    CHECK(fx > 0 && lowered == 0);

    // tilted box with screening, mesh part only: repeatable bits, finite forces
    {
        esp_params pt = make_params(40, 8, 3.0, 3.5, ESP_PREC_FP64);
        pt.alpha = 0.3;
        esp_ctx* ctt = nullptr;
        OK(esp_create(&ctt, &pt));
        esp_system st = sys;
        st.box[3] = 0.25;
        st.box[4] = -0.1;
        st.box[5] = 0.2;
        esp_out ot = out;
        ot.flags = ESP_COMPUTE_MESH | ESP_COMPUTE_ENERGY_VIRIAL;
        OK(esp_compute_forces(ctt, 0, &st, &ot, nullptr));
        std::vector<double> t1(4 * N), t2(4 * N);
        OK(esp_memcpy_dtoh(t1.data(), d_f.p, t1.size() * 8));
        OK(esp_compute_forces(ctt, 1, &st, &ot, nullptr));
        OK(esp_memcpy_dtoh(t2.data(), d_f.p, t2.size() * 8));
        CHECK(all_finite(t1) && std::memcmp(t1.data(), t2.data(), t1.size() * 8) == 0);
        st.box[4] = std::nan("");
        CHECK(esp_compute_forces(ctt, 2, &st, &ot, nullptr) == ESP_ERR_ARG);
        OK(esp_destroy(ctt));
    }

    // batch of this context and a second FP64 one (another mesh and order): each member equals
    // its own evaluation bit for bit; the outputs are cleared first
    esp_params p2 = make_params(36, 6, 2.5, 3.5, ESP_PREC_FP64);
    esp_ctx* ctx2 = nullptr;
    OK(esp_create(&ctx2, &p2));
    esp_out out2 = out;
    out2.force = d_f2.as<double>();
    out2.virial = nullptr;
    out2.ext_energy_virial = nullptr;
    OK(esp_compute_forces(ctx2, 0, &sys, &out2, nullptr));
    std::vector<double> g1(4 * N), b1(4 * N), b2(4 * N);
    OK(esp_memcpy_dtoh(g1.data(), d_f2.p, g1.size() * 8));
    esp_batch* batch = nullptr;
    OK(esp_batch_create(&batch));
    OK(esp_batch_set_options(batch, 1, 0));
    esp_ctx* ctxs[2] = {ctx, ctx2};
    esp_system systems[2] = {sys, sys};
    esp_out outs[2] = {out, out2};
    OK(esp_memset(d_f.p, 0, 4 * N * 8));
    OK(esp_memset(d_f2.p, 0, 4 * N * 8));
    for (int rep = 0; rep < 3; ++rep)
        {
        const int rc = esp_batch_compute(batch, 2, ctxs, systems, outs, nullptr);
        if (rc != ESP_OK)
            std::fprintf(stderr, "esp_batch_compute: %s\n", esp_batch_last_error(batch));
        CHECK(rc == ESP_OK);
        }
    OK(esp_stream_synchronize(nullptr));
    OK(esp_memcpy_dtoh(b1.data(), d_f.p, b1.size() * 8));
    OK(esp_memcpy_dtoh(b2.data(), d_f2.p, b2.size() * 8));
    CHECK(std::memcmp(b1.data(), f1.data(), b1.size() * 8) == 0);
    CHECK(std::memcmp(b2.data(), g1.data(), b2.size() * 8) == 0);
    unsigned long long cap = 0, rep = 0, launches = 0;
    OK(esp_batch_stats(batch, &cap, &rep, &launches));
    CHECK(cap == 1 && rep == 3);
    // a batch mixes no precisions; the mixed-precision context on its own stays near FP64
    esp_params pm = p2;
    pm.precision = ESP_PREC_MIXED;
    esp_ctx* ctx3 = nullptr;
    OK(esp_create(&ctx3, &pm));
    esp_ctx* both[2] = {ctx2, ctx3};
    CHECK(esp_batch_compute(batch, 2, both, systems, outs, nullptr) == ESP_ERR_ARG);
    OK(esp_compute_forces(ctx3, 0, &sys, &out2, nullptr));
    OK(esp_memcpy_dtoh(b2.data(), d_f2.p, b2.size() * 8));
    double d2 = 0, r2 = 0;
    for (unsigned int i = 0; i < N; ++i)
        for (int a = 0; a < 3; ++a)
            {
            d2 += (b2[4 * i + a] - g1[4 * i + a]) * (b2[4 * i + a] - g1[4 * i + a]);
            r2 += g1[4 * i + a] * g1[4 * i + a];
            }
    CHECK(std::sqrt(d2 / r2) < 1e-4);
    OK(esp_destroy(ctx3));
    CHECK(esp_batch_compute(batch, 2, ctxs, nullptr, outs, nullptr) != ESP_OK);
    CHECK(std::strlen(esp_batch_last_error(batch)) > 0);
    OK(esp_batch_destroy(batch));

    // pair paths: allocating, no-allocation and HOOMD's generic pair loop over the evaluator
    OK(esp_pair_forces(ctx, &sys, &out, nullptr));
    OK(esp_memcpy_dtoh(f1.data(), d_f.p, f1.size() * 8));
    OK(esp_pair_forces_noalloc(ctx, &sys, &out, nullptr));
    OK(esp_memcpy_dtoh(f2.data(), d_f.p, f2.size() * 8));
    CHECK(std::memcmp(f1.data(), f2.data(), f1.size() * 8) == 0);
    const double *th = nullptr, *td = nullptr;
    unsigned int nseg = 0;
    OK(esp_get_pair_table(ctx, &th, &td, &nseg));
    esp_pair_param pp;
    std::memset(&pp, 0, sizeof(pp));
    pp.kappa = p.kappa;
    pp.rcut = p.rcut;
    pp.table_host = th;
    pp.table_device = td;
    pp.n_segs = nseg;
    OK(esp_potential_pair_forces(&sys, &pp, p.rcut, 0, &out, nullptr));
    OK(esp_memcpy_dtoh(f2.data(), d_f.p, f2.size() * 8));
    CHECK(all_finite(f2));
    esp_system none = sys;
    none.postype = nullptr;
    CHECK(esp_compute_forces(ctx, 2, &none, &out, nullptr) != ESP_OK);

    // MD: NVE steps with ESP + WCA, then the batched MD entry points
    const double dt = 0.001;
    for (int step = 0; step < 5; ++step)
        {
        OK(esp_integrate_first(N, d_pos.as<double>(), d_vel.as<double>(), d_f.as<double>(), dt, sys.box, nullptr));
        OK(esp_nlist_update(nl, &sys, &rebuilt, nullptr));
        OK(esp_nlist_arrays(nl, &sys.n_neigh, &sys.nlist, &sys.head_list, &nmax));
        OK(esp_compute_forces(ctx, uint64_t(step), &sys, &out, nullptr));
        esp_system inner = sys;
        OK(esp_nlist_inner_arrays(nl, &inner.n_neigh, &inner.nlist, &inner.head_list, &nmax));
        OK(esp_wca_forces(&inner, 1.0, 1.0, &out, nullptr));
        OK(esp_integrate_second(N, d_vel.as<double>(), d_f.as<double>(), dt, 0.0, 1.0, 3, uint64_t(step), nullptr));
        }
    double ke = 0, pe = 0;
    OK(esp_kinetic_potential(N, d_vel.as<double>(), d_f.as<double>(), &ke, &pe));
    CHECK(std::isfinite(ke) && std::isfinite(pe));
    esp_md_batch* mb = nullptr;
    OK(esp_md_batch_create(&mb));
    double* vels[1] = {d_vel.as<double>()};
    esp_nlist* nls[1] = {nl};
    const uint64_t seeds[1] = {9};
    OK(esp_md_batch_set(mb, 1, &sys, vels, &out, nls, seeds));
    OK(esp_md_batch_integrate_first(mb, dt, nullptr));
    unsigned int nreb = 0;
    unsigned char which[1] = {0};
    OK(esp_md_batch_update_nlists(mb, &nreb, which, nullptr));
    OK(esp_md_batch_wca(mb, 1.0, 1.0, nullptr));
    OK(esp_md_batch_integrate_second(mb, dt, 1.0, 1.0, 6, nullptr));
    OK(esp_device_synchronize());
    OK(esp_md_batch_destroy(mb));
    unsigned long long builds = 0;
    OK(esp_nlist_stats(nl, &builds));
    CHECK(builds >= 1);

    OK(esp_destroy(ctx2));
    OK(esp_destroy(ctx));
    OK(esp_nlist_destroy(nl));
    }

int main(int argc, char** argv)
    {
    const bool host_only = argc > 1 && std::strcmp(argv[1], "--host-only") == 0;
    host_tables();
    context_host_paths();
    int ndev = 0;
    if (!host_only && esp_device_count(&ndev) == ESP_OK && ndev > 0)
        {
        gpu_paths();
        std::printf("abi_driver: host and GPU paths, %d failed checks\n", g_fail);
        }
    else
        std::printf("abi_driver: host paths only, %d failed checks\n", g_fail);
    // Leak check now, then leave without the HIP and HSA runtimes' static destructors: their
    // teardown frees runtime memory after the sanitizer's device allocator has been told the
    // runtime unloaded, and the sanitizer's quarantine recycling can then stop on an internal
    // CHECK (sanitizer_allocator_device.h) that is no report about this code.  Leaks are still
    // reported (exit code 23), and everything the driver created has been destroyed above.
    __lsan_do_leak_check();
    std::fflush(stdout);
    std::fflush(stderr);
    _exit(g_fail ? 1 : 0);
    }
