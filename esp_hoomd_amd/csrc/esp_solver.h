// Host orchestration of one ESP solve on one GPU: parameters, tables, device buffers, FFT
// plans and the per-step launch sequence.  Mirrors the stage order of
// ESPForceCompute::computeForces (ESPForceCompute.cc:96-127): one-time initialisation
// (setupMesh, initializeFFT, setupCoeffs/influence, tables), then per step the pair
// correction, assignParticles, forward FFT, influence + energy/virial, inverse FFT,
// interpolateForces and the final reductions.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "esp_kernels.h"
#include "esp_types.h"
#include "pswf_math.h"

namespace esp
{

//! Error carrying a C-ABI status code.
class EspError : public std::runtime_error
    {
    public:
    EspError(int code, const std::string& msg) : std::runtime_error(msg), m_code(code) { }
    int code() const { return m_code; }

    private:
    int m_code;
    };

enum Status
    {
    STATUS_OK = 0,
    STATUS_ARG = -1,
    STATUS_HIP = -2,
    STATUS_OOM = -3,
    STATUS_STATE = -4,
    STATUS_RANGE = -5,
    STATUS_UNSUPPORTED = -6
    };

void check_hip(hipError_t e, const char* what);
//! Compute units of the current device (cached per device).
unsigned int cu_count();

//! RAII device allocation.
class DeviceBuffer
    {
    public:
    DeviceBuffer() = default;
    ~DeviceBuffer();
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    void resize(size_t bytes, bool zero = false);  //!< reallocates only when growing
    //! The same, zeroing with hipMemsetAsync on `s` (the blocking hipMemset waits behind
    //! whatever the device's queues hold).
    void resize_zeroed_async(size_t bytes, hipStream_t s);
    void* get() const { return m_ptr; }
    template<typename T> T* as() const { return static_cast<T*>(m_ptr); }
    size_t bytes() const { return m_bytes; }
    void release();

    private:
    void* m_ptr = nullptr;
    size_t m_bytes = 0;
    };

struct SolverParams
    {
    unsigned int nx = 32, ny = 32, nz = 32;
    unsigned int order = 6;
    double kappa = 0.0;
    double rcut = 0.0;
    double alpha = 0.0;
    unsigned int n_table = 4096;
    int precision = 0;       // 0 FP64, 1 mixed
    int split = 0;           // 0 PSWF, 1 Gaussian
    double window_c = 0.0;   // 0 = automatic
    unsigned int window_degree = 0;  // 0 = automatic
    int pair_lanes = 0;      // lanes per particle in the pair kernel (0 = automatic)
    double k_cut = 0.0;      // reciprocal-space cutoff (0 = automatic, < 0 = none)
    unsigned int tile = 0;   // spreading tile edge (0 = automatic; 2, 4, 8 or 16)
    };

//! Device-side description of one system for one call (caller-owned device pointers).
struct SystemView
    {
    unsigned int N = 0;
    const double* postype = nullptr;
    const double* charge = nullptr;
    unsigned int group_size = 0;
    const unsigned int* group_idx = nullptr;
    double box[6] = {0, 0, 0, 0, 0, 0};
    const unsigned int* n_neigh = nullptr;
    const unsigned int* nlist = nullptr;
    const uint64_t* head_list = nullptr;
    const unsigned int* n_ex = nullptr;
    const unsigned int* ex_list = nullptr;
    unsigned int ex_pitch = 0;
    };

struct OutputView
    {
    double* force = nullptr;
    double* virial = nullptr;
    size_t virial_pitch = 0;
    double* ext_energy_virial = nullptr;  // device double[7] or null
    unsigned int flags = 0;
    };

enum ComputeFlags : unsigned int
    {
    COMPUTE_PAIR = 1u,
    COMPUTE_MESH = 2u,
    COMPUTE_ENERGY_VIRIAL = 4u,
    COMPUTE_ALL = 7u
    };

class EspSolver
    {
    public:
    EspSolver();
    ~EspSolver();

    void set_params(const SolverParams& p);
    const SolverParams& params() const { return m_p; }
    void invalidate()
        {
        m_need_init = true;
        ++m_gen;
        }

    //! Step graphs: compute() captures its whole step (the pair branch on the side stream and
    //! the mesh chain) into a hipGraph once and replays it while the call's arguments (system
    //! and output pointers, sizes, box, flags, stream) and the context's buffers stay the same;
    //! any change captures again.  The replayed kernels read the same device memory, so new
    //! positions or charges in the same arrays are picked up.  Off by default; stage profiling
    //! uses the eager path.
    void set_graph(bool on);
    void graph_stats(unsigned long long& captures, unsigned long long& replays) const
        {
        captures = m_gcaptures;
        replays = m_greplays;
        }

    //! Full ESP evaluation: pair correction + mesh; outputs overwritten.
    void compute(const SystemView& sys, const OutputView& out, hipStream_t s);
    //! The pair correction alone (the kernels of compute with COMPUTE_PAIR) with its packed
    //! neighbour records in caller-provided device memory of `scratch_bytes` bytes (32 per
    //! particle); allocates nothing.  The context must already be initialised (a previous
    //! compute after the last parameter change), else STATUS_STATE; more particles than the
    //! scratch holds give STATUS_RANGE.
    void pair_forces_into(const SystemView& sys, const OutputView& out, hipStream_t s, void* scratch, size_t scratch_bytes);
    //! Prepare this solver for a batched call on `sys` (initialisation, scratch sizing) and
    //! describe its part of the batch in `d`; the batch launches the kernels.
    void describe(const SystemView& sys, const OutputView& out, hipStream_t s, kernel::BatchSys& d);
    //! The set-up the first computation on `sys` would do (validation, tables, influence
    //! function on `s`, scratch sizing), without computing.
    void setup(const SystemView& sys, const OutputView& out, hipStream_t s) { (void)prepare(sys, out, s); }
    bool fp64() const { return m_p.precision == 0; }

    //! Host-side tables (built at initialisation).
    const std::vector<double>& split_screen() const { return m_screen; }
    const std::vector<double>& split_dscreen() const { return m_dscreen; }
    const std::vector<double>& split_r_lo() const { return m_r_lo; }
    const std::vector<double>& split_dr_inv() const { return m_dr_inv; }
    const math::WindowTable& window() const { return m_window; }
    //! Replace the split table by an external one in the PSWF_Coeffs.h format (reduced s).
    void set_split_table(unsigned int n_segs, const double* screen, const double* dscreen, const double* r_lo, const double* dr_inv);
    //! Build host tables now (no device work); used by introspection before the first step.
    void build_host_tables();
    //! The split table as interleaved FP64 records [r_lo, dr_inv, S c0..c4, dS c0..c4], host and
    //! device copies (for the HOOMD-concept pair evaluator); rebuilt on every call.
    void pair_table(const double** host, const double** device, unsigned int* n_segs);
    double self_coeff() const { return m_self_coeff; }
    //! Reciprocal-space cutoff in use: 1.2 kappa by default for the prolate split (beyond the
    //! band limit, where deconvolving the window would only amplify round-off), 0 = none.
    double effective_k_cut() const;
    double background_coeff() const { return m_bg_coeff; }

    //! Charge sums of the last step (synchronises with the device).
    void charge_sums(double& q, double& q2);
    //! Tile edge and tile count chosen for the mesh.
    const MeshGeom& geometry() const { return m_geom; }
    //! Fixed-point scale of the last step's mesh and whether a population bound lowered it
    //! (synchronises; 0 before the first step).
    void fixed_point_scale(float& scale, int& lowered);
    //! Device pointers of the influence tables (half spectrum), for tests.
    const double* device_green() const { return m_G.as<double>(); }
    size_t memory_bytes() const;

    //! Per-stage HIP-event timing (stage boundaries recorded on the compute stream).
    static constexpr int N_STAGES = 12;
    static const char* stage_name(int i);
    void set_profiling(int level);
    //! Totals in ms and call counts per stage since the last reset (synchronises).
    void stage_times(double* ms, unsigned int* calls);
    void reset_stage_times();

    private:
    //! The step after prepare(): pair branch (side stream) and mesh chain on s.
    void enqueue_step(const SystemView& sys, const OutputView& out, unsigned int n, hipStream_t s);
    std::vector<uint64_t> graph_key(const SystemView& sys, const OutputView& out, hipStream_t s) const;
    void release_graph();
    // step-graph replay (set_graph)
    bool m_graph_on = false;
    hipStream_t m_gstream = nullptr;    // capture stream
    hipGraphExec_t m_gexec = nullptr;
    std::vector<uint64_t> m_gkey;
    uint64_t m_gen = 0;                 // bumped whenever a buffer, table or parameter changes
    unsigned long long m_gcaptures = 0, m_greplays = 0;
    void mark(int boundary, hipStream_t s);
    void pair_mark(bool end, hipStream_t s);
    void harvest();
    // side stream for the pair correction, overlapped with the mesh pipeline
    bool m_overlap = true;
    hipStream_t m_side = nullptr;
    hipEvent_t m_ev_fork = nullptr;
    hipEvent_t m_ev_join = nullptr;
    hipEvent_t m_ev_zdone = nullptr;  // the z pass has written the energy/virial partials
    int m_profile = 0;  // 0 off, 1 every stage, 2 the pair stage only
    std::vector<std::vector<hipEvent_t>> m_ev_pool;   // free event sets (N_STAGES + 1 each)
    std::vector<std::vector<hipEvent_t>> m_ev_pending; // recorded, not yet harvested
    std::vector<hipEvent_t>* m_ev_cur = nullptr;
    std::vector<bool> m_ev_marked;
    std::vector<bool> m_ev_full;  // the set holds every stage boundary
    double m_stage_ms[N_STAGES] = {};
    unsigned int m_stage_calls[N_STAGES] = {};

    void initialize(const SystemView& sys, hipStream_t s);
    //! Validate the call, (re)initialise when parameters or the box changed, size the
    //! per-particle scratch; returns the number of mesh-group members.
    unsigned int prepare(const SystemView& sys, const OutputView& out, hipStream_t s);
    kernel::PairArgs pair_args(const SystemView& sys, const OutputView& out) const;
    void build_influence(hipStream_t s);
    void ensure_particles(unsigned int n);
    template<typename R>
    void run_mesh(const SystemView& sys, const OutputView& out, unsigned int n, hipStream_t s, bool want_ev, hipStream_t side);
    //! concurrent: the pass runs on the side stream beside the mesh chain.
    template<typename R> void run_pair(const SystemView& sys, const OutputView& out, hipStream_t s, bool concurrent);

    //! Set-up uploads (tables, twiddles, window, quadrature) go through a pinned staging arena
    //! with hipMemcpyAsync on the set-up's stream, so a set-up does not wait for work already
    //! queued on the device (a server prepares admissions while other requests' steps run).
    //! begin_uploads waits for the previous set-up's copies before the arena is reused.
    void begin_uploads();
    void end_uploads(hipStream_t s);
    template<typename R> void upload_as(DeviceBuffer& buf, const std::vector<double>& v, hipStream_t s);
    unsigned char* m_stage = nullptr;
    size_t m_stage_cap = 0, m_stage_used = 0;
    hipEvent_t m_stage_ev = nullptr;
    hipStream_t m_stage_stream = nullptr;
    bool m_stage_recorded = false;
    std::vector<unsigned char*> m_stage_retired;  // outgrown blocks, freed by the next set-up

    SolverParams m_p;
    bool m_need_init = true;
    bool m_tables_built = false;
    bool m_external_table = false;
    double m_box[6] = {0, 0, 0, 0, 0, 0};  // Lx, Ly, Lz, xy, xz, yz of the tables' box
    void set_box_geometry(const double* box);
    MeshGeom m_geom {};
    FFTPlan m_plan[3] {};
    unsigned int m_n_tiles = 0;
    unsigned int m_D = 0;
    double m_self_coeff = 0.0;
    double m_bg_coeff = 0.0;

    std::vector<double> m_screen, m_dscreen, m_r_lo, m_dr_inv;
    math::WindowTable m_window;
    std::vector<double> m_qx, m_qw;

    // device
    static constexpr unsigned int COMPACT_SEGS = 64;
    DeviceBuffer m_split_tab;   // [n_segs][12] R
    DeviceBuffer m_compact;     // [COMPACT_SEGS][compact_deg() + 1] R
    unsigned int compact_deg() const { return m_p.precision == 0 ? kernel::pair_compact_deg<double>() : kernel::pair_compact_deg<float>(); }
    DeviceBuffer m_wcoef;       // [P][D+1] R
    DeviceBuffer m_tw[3];       // twiddles per axis, complex R
    DeviceBuffer m_G, m_Gv;     // half spectrum FP64
    DeviceBuffer m_qxd, m_qwd, m_w2[3];
    DeviceBuffer m_mesh;        // spread accumulator (FP64 or int64 fixed point)
    DeviceBuffer m_spec;        // half spectrum complex R
    DeviceBuffer m_phi;         // potential mesh R
    DeviceBuffer m_tile_count, m_tile_start;
    DeviceBuffer m_key, m_slot, m_rec, m_sidx;
    DeviceBuffer m_pq;          // packed (x, y, z, q) per particle for the pair pass
    std::vector<double> m_pair_rec;  // interleaved FP64 split table for the pair evaluator
    DeviceBuffer m_pair_rec_dev;
    DeviceBuffer m_qpart, m_evpart, m_qsum, m_ev_internal;
    DeviceBuffer m_fx;          // the mesh's fixed-point scale, its inverse, lowered flag (float[4])
    float m_w3max = 1.0f;       // largest |w_x w_y w_z| of the window
    //! The scan's rule for the fixed-point scale: int32 LDS cells for FP32 taps, int64 cells
    //! (over the tile boxes that overlap a cell) for FP64 taps.
    FixedPointRule fixed_point_rule() const
        {
        return FixedPointRule {m_w3max, tile_box_cover(m_geom.T, int(m_p.order)), m_p.precision == 0 ? 1 : 0};
        }
    unsigned int m_cap_particles = 0;
    };

//! Mixed-radix plan for length n using radices 4, 2, 3, 5; throws for other factors.
FFTPlan make_fft_plan(int n);

    } // namespace esp
