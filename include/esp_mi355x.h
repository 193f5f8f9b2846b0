/*
 * esp_mi355x.h -- C ABI of the MI355X-native ESP (Ewald summation with prolates) solver.
 *
 * Plain C types only: device pointers, sizes and status codes.  No C++ exceptions cross
 * this boundary.  A context is single threaded and enqueues all work on the caller's
 * stream; distinct contexts may run concurrently.  The caller owns the system and output
 * buffers; the library owns meshes, tables and plans.
 *
 * Units and layouts are those of the reference plugin:
 *   Coulomb kernel 1/(4 pi r)                         (ESPForceCompute.cc:36)
 *   postype Scalar4[N] = (x, y, z, type), box centred on the origin, charge Scalar[N]
 *   force   Scalar4[N] = (fx, fy, fz, per-particle energy)
 *   virial  Scalar[6 * pitch] SoA xx, xy, xz, yy, yz, zz  (ESPForceComputeGPU.cu:453-458)
 *   mesh index (cz * Ny + cy) * Nx + cx                   (ESPForceComputeGPU.cu:267)
 *   neighbour list: full list, nlist[head_list[i] + k] for k < n_neigh[i] (HOOMD layout)
 *   exclusions: ex_list[i * ex_pitch + k] for k < n_ex[i]  (Index2D(ex, i), ESPForceComputeGPU.cu:418)
 * Scalar is double.
 */
#ifndef ESP_MI355X_H_
#define ESP_MI355X_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define ESP_OK 0
#define ESP_ERR_ARG (-1)
#define ESP_ERR_HIP (-2)
#define ESP_ERR_OOM (-3)
#define ESP_ERR_STATE (-4)
#define ESP_ERR_RANGE (-5)
#define ESP_ERR_UNSUPPORTED (-6)

/* precision modes */
#define ESP_PREC_FP64 0   /* every stage in FP64; the mesh sums FP64 stencil products in int64
                             fixed point, so results are bit-reproducible (run to run, any
                             particle order or tiling, alone or in a batch) */
#define ESP_PREC_MIXED 1  /* FP32 spread/FFT/gather/pair, fixed-point mesh accumulation,
                             FP64 influence function, self/k=0 terms and reductions */

/* split kinds */
#define ESP_SPLIT_PSWF 0      /* the ESP prolate split (compactly supported short range) */
#define ESP_SPLIT_GAUSSIAN 1  /* erfc split (textbook Ewald / SPME), for validation */

/* compute flags (esp_out.flags; 0 means all) */
#define ESP_COMPUTE_PAIR 1u
#define ESP_COMPUTE_MESH 2u
#define ESP_COMPUTE_ENERGY_VIRIAL 4u
#define ESP_COMPUTE_ALL 7u

typedef struct esp_ctx esp_ctx;
typedef void* esp_stream_t; /* a hipStream_t; NULL is the default stream */

/* Parameters of ESPForceCompute::setParams(nx, ny, nz, order, kappa, rcut, alpha, n_table)
 * (ESPForceCompute.h:99-106) plus the MI355X-specific options.  kappa is the splitting
 * parameter in 1/length: for the PSWF split the prolate bandwidth is c = kappa * rcut and
 * the long-range part is band limited near |k| = kappa. */
typedef struct esp_params {
    unsigned int nx, ny, nz;       /* mesh resolution; each a product of 2, 3, 5 */
    unsigned int order;            /* stencil order P, 2..16 */
    double kappa;                  /* splitting parameter [1/length] */
    double rcut;                   /* real-space cutoff r_c */
    double alpha;                  /* Debye screening >= 0: the influence function is multiplied by
                                      exp(-alpha |k|) (ESPForceComputeGPU.cu:199); 0 = none */
    unsigned int n_table;          /* split-table segments (PSWF_N_SEGS = 4096 default) */
    int precision;                 /* ESP_PREC_* */
    int split;                     /* ESP_SPLIT_* */
    double window_c;               /* window prolate bandwidth; 0 = automatic */
    unsigned int window_degree;    /* window polynomial degree; 0 = automatic */
    int pair_lanes;                /* lanes per particle in the pair kernel; 0 = automatic */
    double k_cut;                  /* spherical reciprocal-space cutoff: the influence function is 0
                                      for |k| > k_cut.  0 = automatic (1.2 kappa for the PSWF split,
                                      none for the Gaussian split); < 0 = none */
    unsigned int tile;             /* spreading/gathering tile edge in mesh cells, 2 to 16, whose
                                      (tile + order - 1)^3 LDS box fits 64 KiB (8-byte cells in
                                      FP64, 4-byte in mixed); 0 = automatic */
} esp_params;

/* One system: device pointers owned by the caller (HOOMD's ParticleData, NeighborList and
 * ParticleGroup arrays, ESPForceComputeGPU.cc:221-224, 397-402). */
typedef struct esp_system {
    unsigned int N;                 /* particles in postype/charge/force */
    const double* postype;          /* Scalar4[N] */
    const double* charge;           /* Scalar[N] */
    unsigned int group_size;        /* members of the mesh group (ignored if group_idx NULL) */
    const unsigned int* group_idx;  /* group member indices, or NULL for all N */
    double box[6];                  /* HOOMD BoxDim: Lx, Ly, Lz and tilt factors xy, xz, yz (lattice vectors
                                       (Lx,0,0), (xy Ly,Ly,0), (xz Lz,yz Lz,Lz)); spreading and gathering use
                                       makeFraction (ESPForceComputeGPU.cu:228, 320), pairs HOOMD's minImage,
                                       so in a tilted box rcut must stay below half the box's smallest
                                       perpendicular width */
    const unsigned int* n_neigh;    /* full neighbour list, or NULL for no pair part */
    const unsigned int* nlist;
    const uint64_t* head_list;
    const unsigned int* n_ex;       /* exclusions, or NULL */
    const unsigned int* ex_list;
    unsigned int ex_pitch;
} esp_system;

/* Outputs, overwritten each call (ForceCompute m_force / m_virial / external energy). */
typedef struct esp_out {
    double* force;                  /* Scalar4[N] */
    double* virial;                 /* Scalar[6 * virial_pitch], or NULL */
    size_t virial_pitch;            /* >= N */
    double* ext_energy_virial;      /* device double[7]: E_ext, W_xx, W_xy, W_xz, W_yy, W_yz, W_zz; or NULL */
    unsigned int flags;             /* ESP_COMPUTE_*; 0 = all */
} esp_out;

/* ESPTableEntry (ESPForceCompute.h:78-88): introspection table of L(r) and -dL/dr. */
typedef struct esp_table_entry {
    double potential_coeffs[4];
    double force_coeffs[4];
    double potential_coeff4;
    double force_coeff4;
    double r_lo;
    double dr_inv;
} esp_table_entry;

/* Introspection (getters of ESPForceCompute.h:108-118). */
typedef struct esp_info {
    unsigned int nx, ny, nz;
    unsigned int order;
    double kappa, rcut, alpha;
    unsigned int n_table;
    int precision, split;
    unsigned int window_degree;
    double window_c;
    double self_coeff;              /* lim (1 - S(r/rc)) / r */
    double background_coeff;        /* lim (1 - F(k)) / k^2 */
    unsigned int tile;              /* spreading tile edge (0 before the first step) */
    size_t device_bytes;            /* device memory held by the context */
    double k_cut;                   /* effective reciprocal-space cutoff (0 = none) */
} esp_info;

/* ---- lifetime and parameters --------------------------------------------------------- */
/* ESPForceCompute ctor + setParams (ESPForceCompute.cc:53-89; pybind ctor :454-467) */
int esp_create(esp_ctx** ctx, const esp_params* params);
/* ESPForceCompute::setParams (ESPForceCompute.cc:71-89); takes effect on the next compute */
int esp_set_params(esp_ctx* ctx, const esp_params* params);
/* ESPForceCompute::invalidate (ESPForceCompute.h:120) */
int esp_invalidate(esp_ctx* ctx);
/* destructor (ESPForceCompute.cc:69) */
int esp_destroy(esp_ctx* ctx);
/* last error message of this context (or of the last failed esp_create when ctx is NULL) */
const char* esp_last_error(const esp_ctx* ctx);

/* ---- compute --------------------------------------------------------------------------- */
/* ESPForceCompute::computeForces (ESPForceCompute.cc:96-127) on the GPU: pair correction
 * (EvaluatorPairPSWF over the neighbour list, EvaluatorPairPSWF.h:274-295) + exclusion
 * correction (ESPForceComputeGPU.cu:391-471) + mesh part (.cu:206-373) + energy/virial. */
int esp_compute_forces(esp_ctx* ctx, uint64_t timestep, const esp_system* sys, esp_out* out, esp_stream_t stream);
/* Step graphs (no counterpart in the reference, which launches every kernel per step): with
 * enable = 1, esp_compute_forces captures its whole step (pair branch and mesh chain) into a
 * hipGraph on the first call and replays it on later calls whose arguments (system and output
 * pointers, sizes, box, flags, stream) and parameters are unchanged, cutting the per-kernel
 * launch gaps; any change captures again.  New contents of the same arrays (positions,
 * charges, neighbour lists) are picked up.  Mesh-only calls (flags without ESP_COMPUTE_PAIR)
 * and profiled calls stay eager.  Default 0. */
int esp_set_graph(esp_ctx* ctx, int enable);
int esp_get_graph_stats(esp_ctx* ctx, unsigned long long* captures, unsigned long long* replays);
/* The set-up esp_compute_forces does lazily on its first call after esp_create,
 * esp_set_params or a box change (ESPForceCompute::computeForces initialises when
 * m_need_initialize is set, ESPForceCompute.cc:98-107): arguments validated, tables and
 * window built, the influence function enqueued on `stream`, scratch sized -- nothing
 * computed.  For callers that want that work off a step's critical path, e.g. a server
 * admitting a request while other requests' steps run.  Errors are those of
 * esp_compute_forces. */
int esp_prepare(esp_ctx* ctx, const esp_system* sys, const esp_out* out, esp_stream_t stream);
/* the PotentialPair<EvaluatorPairPSWF> path alone (module.cc:73,79): pair + exclusions */
int esp_pair_forces(esp_ctx* ctx, const esp_system* sys, esp_out* out, esp_stream_t stream);

/* No counterpart in the reference: esp_pair_forces for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  The packed neighbour records
 * live in a static device array of the library (32 bytes per particle, ESP_NOALLOC_MAX_PARTICLES
 * particles, 8 MiB held on every GPU the library runs on until the process ends); nothing is
 * allocated.  Same kernels and results as esp_pair_forces.  The context must have been
 * initialised by one allocating call (esp_compute_forces or esp_pair_forces) after esp_create /
 * esp_set_params, else ESP_ERR_STATE; N above the maximum gives ESP_ERR_RANGE.  All calls on a
 * GPU share its one scratch array: calls that overlap in time on the same GPU (different
 * streams or threads) race. */
#define ESP_NOALLOC_MAX_PARTICLES 262144u
int esp_pair_forces_noalloc(esp_ctx* ctx, const esp_system* sys, esp_out* out, esp_stream_t stream);

/* ---- the pair evaluator in HOOMD's concept (include/EvaluatorPairPSWF.h) ------------------
 * esp_pair_param mirrors EvaluatorPairPSWF::param_type (EvaluatorPairPSWF.h:180-225 of the
 * reference, plus the split table the evaluator reads): kappa, alpha, rcut, V_cut, the table's
 * host and device addresses and its segment count. */
typedef struct esp_pair_param {
    double kappa, alpha, rcut, V_cut;
    const double* table_host;
    const double* table_device;
    unsigned int n_segs;
    unsigned int pad[3];  /* 64 bytes, as the 16-byte aligned param_type */
} esp_pair_param;
/* The context's split table for the evaluator: n_segs records of 12 doubles (r_lo, dr_inv,
 * S c0..c4, dS/ds c0..c4, reduced s = r / rc), one host and one device copy owned by the
 * context (getTablePtr, ESPForceCompute.cc:129-132, returns the host address).  Valid until the
 * context's parameters or table change or the context is destroyed; call again after
 * esp_set_params / esp_set_split_table. */
int esp_get_pair_table(esp_ctx* ctx, const double** host_table, const double** device_table, unsigned int* n_segs);
/* PotentialPairGPU<EvaluatorPairPSWF>::computeForces (module.cc:79): HOOMD's generic pair loop,
 * one thread per particle over the full neighbour list, the evaluator of EvaluatorPairPSWF.h for
 * every pair within r_cut; force, per-particle energy (half of each pair) and virial (half of
 * each pair) overwritten.  Excluded pairs must be absent from the list, as HOOMD builds it. */
int esp_potential_pair_forces(const esp_system* sys, const esp_pair_param* param, double r_cut, int energy_shift, esp_out* out,
                              esp_stream_t stream);

/* ---- batched solves --------------------------------------------------------------------
 * Many independent systems, one launch per stage (no counterpart in the reference, which
 * evaluates one system per ESPForceCompute).  System i is evaluated with the parameters,
 * tables and scratch of ctxs[i] exactly as esp_compute_forces(ctxs[i], ...) would, and its
 * outputs are overwritten.  All contexts of a call share one precision, a context appears at
 * most once per call, and stencil orders are 4, 6, 8, 10 or 12.  The launch sequence is
 * captured into a hipGraph and replayed while the batch keeps its shape (same contexts,
 * particle counts, meshes and flags); new device pointers are picked up without a new
 * capture.  A batch is single threaded; its contexts must not be used concurrently. */
typedef struct esp_batch esp_batch;
int esp_batch_create(esp_batch** batch);
int esp_batch_destroy(esp_batch* batch);
const char* esp_batch_last_error(const esp_batch* batch);
/* use_graph: capture and replay (default 1); max_grid: blocks per batched launch that walk
 * the work items with a grid stride (default 2048; 0 = one block per item) */
int esp_batch_set_options(esp_batch* batch, int use_graph, unsigned int max_grid);
int esp_batch_compute(esp_batch* batch, unsigned int n, esp_ctx* const* ctxs, const esp_system* systems, const esp_out* outs,
                      esp_stream_t stream);
int esp_batch_stats(esp_batch* batch, unsigned long long* captures, unsigned long long* replays, unsigned long long* launches);

/* ---- introspection ------------------------------------------------------------------- */
int esp_get_info(esp_ctx* ctx, esp_info* info);
/* getQSum / getQ2Sum (ESPForceCompute.h:113-114); synchronises with the device */
int esp_get_charge_sums(esp_ctx* ctx, double* q_sum, double* q2_sum);
/* The fixed-point scale of the last step's mesh accumulation: mixed precision 2^23 / 2^ceil(log2
 * max|q|), FP64 2^50 / 2^ceil(log2 max|q|); lower when the most populated tile could overflow
 * its accumulators (int32 LDS cells; int64 cells summed over the overlapping tile boxes), then
 * *lowered = 1; 0 before the first step.  Synchronises.  No counterpart in the reference,
 * which sums the mesh with FP64 atomics. */
int esp_get_fixed_point_scale(esp_ctx* ctx, float* scale, int* lowered);
/* getTableSize / getTableEntry (ESPForceCompute.cc:129-141, 336-372) */
int esp_get_table_size(esp_ctx* ctx, unsigned int* n);
int esp_get_table_entry(esp_ctx* ctx, unsigned int i, esp_table_entry* entry);
/* split table in the PSWF_Coeffs.h format (PSWF_Coeffs.h:24-9599), reduced s = r / rc:
 * screen[n][5], dscreen[n][5] (dS/ds), r_lo[n], dr_inv[n]; pass NULL arrays to query n */
int esp_get_split_table(esp_ctx* ctx, unsigned int* n_segs, double* screen, double* dscreen, double* r_lo, double* dr_inv);
/* replace the split table (e.g. the arrays of a PSWF_Coeffs.h header) */
int esp_set_split_table(esp_ctx* ctx, unsigned int n_segs, const double* screen, const double* dscreen, const double* r_lo,
                        const double* dr_inv);
/* spreading window: P taps x (D+1) monomial coefficients in z = 2t - 1 (replaces
 * PSWF_STENCIL_COEFFS / _X_LO / _DX_INV, PSWF_Coeffs.h:9606-9624) */
int esp_get_window(esp_ctx* ctx, unsigned int* P, unsigned int* D, double* coef);

/* ---- tracing -------------------------------------------------------------------------- */
/* Per-stage HIP-event timing on the compute stream (no counterpart in the reference, which
 * only brackets kernels with HOOMD Autotuner begin/end, ESPForceComputeGPU.cc:230-250). */
#define ESP_N_STAGES 12
/* enable: 0 off, 1 every stage (an event pair around each), 2 the pair stage only (events on
 * the pair kernel's stream; the step's other kernels run unmarked) */
int esp_set_profiling(esp_ctx* ctx, int enable);
/* totals in ms and call counts for the ESP_N_STAGES stages since the last reset; synchronises */
int esp_get_stage_times(esp_ctx* ctx, double* ms, unsigned int* calls);
int esp_reset_stage_times(esp_ctx* ctx);
const char* esp_stage_name(int i);

/* ---- host-only table mathematics (no GPU needed) ------------------------------------- */
/* table generator (replaces generate_pswf_table.py:540-595 with a valid split, S(0) = 1) */
int esp_build_split_table(int split, double kappa, double rcut, unsigned int n_segs, double* screen, double* dscreen, double* r_lo,
                          double* dr_inv);
int esp_split_eval(int split, double kappa, double rcut, int n, const double* s, double* S, double* dSds);
int esp_split_fourier(int split, double kappa, double rcut, int n, const double* k, double* F, double* dF);
int esp_split_constants(int split, double kappa, double rcut, double* self_coeff, double* background_coeff);
/* prolate psi_0^c at n points: psi[i] = psi_0^c(x[i]), integral[i] = int_0^x[i] psi_0^c(t) dt */
int esp_prolate_eval(double c, int n, const double* x, double* psi, double* integral);
int esp_build_window(unsigned int P, unsigned int D, double window_c, double* coef);
int esp_window_fourier(unsigned int P, unsigned int D, const double* coef, int n, const double* theta, double* what);
double esp_default_window_c(unsigned int P, double sigma);
unsigned int esp_default_window_degree(unsigned int P);

/* ---- MD support: the roles HOOMD plays around the plugin ---------------------------------
 * Neighbour list (HOOMD NeighborList consumed by ESPForceCompute, ESPForceCompute.h:39,177 and
 * by PotentialPair<EvaluatorPairPSWF>, module.cc:73,79): cell-list build in the full-list
 * layout, rebuilt when any particle moved more than r_buff / 2. */
typedef struct esp_nlist esp_nlist;
int esp_nlist_create(esp_nlist** nl, double r_cut, double r_buff);
int esp_nlist_destroy(esp_nlist* nl);
const char* esp_nlist_last_error(const esp_nlist* nl);
/* build now (uses sys->postype, N, box and the exclusion list; synchronises the stream) */
int esp_nlist_build(esp_nlist* nl, const esp_system* sys, esp_stream_t stream);
/* rebuild if needed; *rebuilt tells whether it did (synchronises the stream) */
int esp_nlist_update(esp_nlist* nl, const esp_system* sys, int* rebuilt, esp_stream_t stream);
/* device arrays of the current list (HOOMD layout) */
int esp_nlist_arrays(esp_nlist* nl, const unsigned int** n_neigh, const unsigned int** nlist, const uint64_t** head_list,
                     unsigned int* nmax);
int esp_nlist_stats(esp_nlist* nl, unsigned long long* builds);
/* The particle indices in cell order of the last build (members of cell 0, then cell 1, ...;
 * ascending within a cell): a permutation of [0, N) for esp_gather_rows.  HOOMD keeps its
 * particles in space-filling-curve order with its SFCPack updater; a caller without one sorts
 * with this (no counterpart in the reference).  ESP_ERR_STATE before the first build. */
int esp_nlist_cell_order(esp_nlist* nl, const unsigned int** members, unsigned int* n);
/* dst row i = src row order[i] for n rows of row_bytes bytes (a multiple of 4) on `stream`:
 * permutes per-particle arrays into the order above (no counterpart in the reference).  Rows
 * whose order entry is >= n are left unwritten. */
int esp_gather_rows(unsigned int n, unsigned int row_bytes, const unsigned int* order, const void* src, void* dst,
                    esp_stream_t stream);
/* Inner list (no counterpart in the reference, where a HOOMD list is built at the largest cutoff
 * of all pair forces and each force walks all of it): with r_inner > 0 (and below r_cut) every
 * build also keeps, in its own rows, the pairs within r_inner + r_buff, which is a valid list for
 * any pair force with cutoff <= r_inner under the same rebuild rule (the batched WCA uses it);
 * 0 turns it off.  Takes effect at the next build (the next update rebuilds). */
int esp_nlist_set_inner(esp_nlist* nl, double r_inner);
/* Change the list's cutoff (the next update rebuilds); its arrays are kept and reused by the
 * builds that fit them (no counterpart in the reference: lets a server hand a retired request's
 * list to a new one without reallocating) */
int esp_nlist_set_cutoff(esp_nlist* nl, double r_cut);
/* Row capacities (neighbours per particle) of the next builds, full and inner list; nmax = 0
 * returns to the automatic sizing from the density (1.5x the mean count plus 32).  A build that
 * overflows its rows raises them and is repeated (esp_nlist_build / _update) or reported
 * (esp_md_batch_finish_rebuilds).  No counterpart in the reference. */
int esp_nlist_set_capacity(esp_nlist* nl, unsigned int nmax, unsigned int nmax_inner);
/* device arrays of the current inner list (HOOMD layout); ESP_ERR_STATE if there is none */
int esp_nlist_inner_arrays(esp_nlist* nl, const unsigned int** n_neigh, const unsigned int** nlist, const uint64_t** head_list,
                           unsigned int* nmax);
/* WCA repulsion (LJ truncated at 2^(1/6) sigma, shifted) added to out->force / out->virial */
int esp_wca_forces(const esp_system* sys, double sigma, double epsilon, esp_out* out, esp_stream_t stream);
/* velocity Verlet: vel = (vx, vy, vz, mass) per particle; box = (Lx, Ly, Lz) */
int esp_integrate_first(unsigned int N, double* postype, double* vel, const double* force, double dt, const double* box,
                        esp_stream_t stream);
/* second half; gamma > 0 adds a Langevin bath at temperature kT (counter-based RNG: seed, step) */
int esp_integrate_second(unsigned int N, double* vel, const double* force, double dt, double gamma, double kT, uint64_t seed,
                         uint64_t step, esp_stream_t stream);
/* total kinetic energy and sum of per-particle energies (force.w); synchronises */
int esp_kinetic_potential(unsigned int N, const double* vel, const double* force, double* ke, double* pe);

/* Batched MD: the operations above for many independent systems, one launch each (no
 * counterpart in the reference; used by the serving loop).  esp_md_batch_set records the
 * members: systems[i] (N, postype, box), vel[i], outs[i] (force, virial, pitch), nls[i] (the
 * system's neighbour list or NULL: no WCA, no rebuilds) and seeds[i] (Langevin stream, or
 * NULL for 0).  esp_md_batch_update_nlists rebuilds the lists of members that moved more
 * than r_buff / 2 (one displacement launch for the batch and one host synchronisation);
 * which[i] (optional, host) tells whose list was rebuilt, so callers refresh the list
 * pointers they pass elsewhere (esp_nlist_arrays). */
typedef struct esp_md_batch esp_md_batch;
int esp_md_batch_create(esp_md_batch** batch);
int esp_md_batch_destroy(esp_md_batch* batch);
const char* esp_md_batch_last_error(const esp_md_batch* batch);
int esp_md_batch_set(esp_md_batch* batch, unsigned int n, const esp_system* systems, double* const* vel, const esp_out* outs,
                     esp_nlist* const* nls, const uint64_t* seeds);
/* esp_md_batch_set with the descriptor uploads enqueued on `stream` from pinned staging (no
 * wait); the caller synchronises with the stream before the next set of the same batch */
int esp_md_batch_set_async(esp_md_batch* batch, unsigned int n, const esp_system* systems, double* const* vel, const esp_out* outs,
                           esp_nlist* const* nls, const uint64_t* seeds, esp_stream_t stream);
/* Members with hold[i] != 0 are skipped by the next integrate calls until the next set or
 * set_hold (a member whose first forces are computed in this step); the flags are uploaded on
 * `stream` (synchronise with it before the next set_hold); hold = NULL or all zero clears */
int esp_md_batch_set_hold(esp_md_batch* batch, const unsigned char* hold, esp_stream_t stream);
int esp_md_batch_integrate_first(esp_md_batch* batch, double dt, esp_stream_t stream);
int esp_md_batch_integrate_second(esp_md_batch* batch, double dt, double gamma, double kT, uint64_t step, esp_stream_t stream);
int esp_md_batch_wca(esp_md_batch* batch, double sigma, double epsilon, esp_stream_t stream);
int esp_md_batch_update_nlists(esp_md_batch* batch, unsigned int* rebuilt, unsigned char* which, esp_stream_t stream);
/* List upkeep without a wait inside the step (no counterpart in the reference).  At the end of
 * a step, esp_md_batch_predict(dt) enqueues, per member, the largest squared distance between
 * the position the next esp_md_batch_integrate_first(dt) will produce (same arithmetic, same
 * bits) and the position at the last build, copied into pinned host memory.  Once the caller
 * has synchronised with that step, esp_md_batch_read_predicted sets need[i] = 1 for the
 * members whose list the next step must rebuild (the esp_md_batch_update_nlists criterion,
 * so the decisions are identical).  In the next step, after integrate_first,
 * esp_md_batch_rebuild_async(which = need) enqueues those builds and records their positions
 * without waiting; moved[i] (optional) = 1 where the list arrays were reallocated (callers
 * refresh the pointers they pass elsewhere).  After synchronising with that step,
 * esp_md_batch_finish_rebuilds reports in overflowed[i] the lists whose rows were too short:
 * their capacities are raised and the list marked unbuilt, the member's forces of that step are
 * invalid and the caller repeats its step (e.g. from a copy of its state taken before it). */
int esp_md_batch_predict(esp_md_batch* batch, double dt, esp_stream_t stream);
int esp_md_batch_read_predicted(esp_md_batch* batch, unsigned char* need);
int esp_md_batch_rebuild_async(esp_md_batch* batch, const unsigned char* which, unsigned char* moved, esp_stream_t stream);
int esp_md_batch_finish_rebuilds(esp_md_batch* batch, unsigned char* overflowed, unsigned int* n_overflowed);

/* ---- device utilities (plumbing for callers without a device allocator) ---------------- */
int esp_device_count(int* n);
int esp_set_device(int device);
int esp_device_malloc(void** ptr, size_t bytes);
int esp_device_free(void* ptr);
int esp_memcpy_htod(void* dst, const void* src, size_t bytes);
int esp_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int esp_memcpy_dtod(void* dst, const void* src, size_t bytes, esp_stream_t stream);
/* host-to-device copy enqueued on a stream: from pinned memory (esp_host_malloc_pinned) the call
 * returns at once, and the source must stay unchanged until the stream has passed the copy */
int esp_memcpy_htod_async(void* dst, const void* src, size_t bytes, esp_stream_t stream);
int esp_host_malloc_pinned(void** ptr, size_t bytes);
int esp_host_free_pinned(void* ptr);
int esp_memset(void* dst, int value, size_t bytes);
/* copy n 32-byte records with one vector load and one vector store per record (a plain
   streaming kernel of known traffic: 64 bytes of HBM per record) */
int esp_copy_records(void* dst, const void* src, unsigned int n_records, esp_stream_t stream);
int esp_stream_create(esp_stream_t* stream);
int esp_stream_destroy(esp_stream_t stream);
int esp_stream_synchronize(esp_stream_t stream);
int esp_device_synchronize(void);
/* HIP events for timing on a given stream */
int esp_event_create(void** ev);
int esp_event_destroy(void* ev);
int esp_event_record(void* ev, esp_stream_t stream);
int esp_event_elapsed_ms(void* start, void* stop, float* ms);
const char* esp_version(void);

#ifdef __cplusplus
}
#endif

#endif /* ESP_MI355X_H_ */
