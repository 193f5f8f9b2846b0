// Batched ESP solves: many independent systems (each with its own EspSolver holding its
// parameters, tables and scratch) evaluated by one launch per stage.  The launch sequence
// of a batch is captured once into a hipGraph and replayed while the batch keeps its shape
// (the same solvers, particle counts, meshes and flags); changed device pointers only
// refresh the per-system descriptor table that the kernels read.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "esp_kernels.h"
#include "esp_solver.h"

namespace esp
{

class EspBatch
    {
    public:
    EspBatch();
    ~EspBatch();
    EspBatch(const EspBatch&) = delete;
    EspBatch& operator=(const EspBatch&) = delete;

    //! Evaluate n systems; solvers[i] holds the parameters of system i (all of one precision).
    //! Outputs are overwritten exactly as EspSolver::compute does.
    void compute(unsigned int n, EspSolver* const* solvers, const SystemView* sys, const OutputView* out, hipStream_t s);

    //! use_graph: capture/replay the launch sequence; max_grid: blocks per batched launch
    //! (0 = one block per work item).
    void set_options(bool use_graph, unsigned int max_grid);

    uint64_t captures() const { return m_captures; }
    uint64_t replays() const { return m_replays; }
    uint64_t launches() const { return m_launches; }

    private:
    struct Lists
        {
        kernel::BatchList pack, pair, bin, scan, scatter, tiles, fft_x, fft_y, fft_z, fft_xy, fin;
        };
    void build_plan(bool fp64, hipStream_t s);
    void enqueue(hipStream_t s);
    template<typename R> void enqueue_stages(hipStream_t s);
    void release_graph();

    bool m_use_graph = true;
    unsigned int m_max_grid = 2048;
    bool m_fp64 = false;
    std::vector<kernel::BatchSys> m_desc;     // host copy of the current descriptors
    std::vector<kernel::BatchSys> m_uploaded; // what the device table holds
    std::vector<uint64_t> m_shape;            // shape key of the current plan
    kernel::BatchSys* m_staging = nullptr;    // pinned upload staging
    size_t m_staging_cap = 0;
    kernel::BatchItem* m_items_staging = nullptr;  // pinned staging of the work lists
    size_t m_items_staging_cap = 0;
    hipEvent_t m_upload_done = nullptr;
    DeviceBuffer m_desc_dev;
    DeviceBuffer m_items_dev;
    Lists m_l {};
    size_t m_lds_stencil = 0, m_lds_fx = 0, m_lds_fy = 0, m_lds_fz = 0, m_lds_fxy = 0, m_lds_pair = 0;
    bool m_any_mesh = false;
    bool m_any_tilted = false;  // some system of the plan has a tilted box

    hipStream_t m_cap = nullptr;   // capture stream
    hipStream_t m_side = nullptr;  // pair branch
    hipEvent_t m_fork = nullptr, m_join = nullptr, m_done = nullptr;
    hipGraph_t m_graph = nullptr;
    hipGraphExec_t m_exec = nullptr;
    uint64_t m_captures = 0, m_replays = 0, m_launches = 0;
    };

    } // namespace esp
