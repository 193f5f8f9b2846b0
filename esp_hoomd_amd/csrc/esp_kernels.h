// Kernel-launcher layer: host-callable functions returning hipError_t, one per device stage
// (the role of ESPForceComputeGPU.cuh:49-144 in the reference).  Every launcher enqueues on the
// given stream and performs no synchronisation.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "esp_types.h"

namespace esp
{
namespace kernel
{

// ---- binning / sorting -----------------------------------------------------
hipError_t bin_particles(unsigned int n,
                         const unsigned int* group,
                         const double* postype,
                         const double* charge,
                         const MeshGeom& g,
                         int P,
                         unsigned int* tile_count,
                         unsigned int* key,
                         unsigned int* slot,
                         double* qpart,
                         hipStream_t s);
//! Exclusive scan of the tile counts (which it clears); with fx it also writes the step's
//! fixed-point scale from the largest tile count and |q| (qpart: 3 per binning block).
hipError_t scan_tiles(unsigned int n_tiles, unsigned int* count, unsigned int* start, int n_q, const double* qpart, FixedPointRule rule,
                      float* fx, hipStream_t s);
template<typename R>
hipError_t scatter_particles(unsigned int n,
                             const unsigned int* group,
                             const double* postype,
                             const double* charge,
                             const MeshGeom& g,
                             int P,
                             const unsigned int* key,
                             const unsigned int* slot,
                             const unsigned int* start,
                             void* rec,
                             SortedIdx* sidx,
                             hipStream_t s);

// ---- spread / gather ------------------------------------------------------
size_t spread_lds_bytes(const MeshGeom& g, int P, int D, bool fp64);
size_t gather_lds_bytes(const MeshGeom& g, int P, int D, bool fp64);
template<typename R>
hipError_t spread(const MeshGeom& g, int P, int D, unsigned int n_tiles, const unsigned int* start, const void* rec,
                  const SortedIdx* sidx, const void* wcoef, void* mesh, const float* fx, hipStream_t s);
template<typename R>
hipError_t gather(const MeshGeom& g, int P, int D, unsigned int n_tiles, const unsigned int* start, const void* rec,
                  const SortedIdx* sidx, const void* wcoef, const void* phi, double* force, hipStream_t s);

// ---- FFT passes -------------------------------------------------------------
int fft_x_pairs_per_block(int nx, bool fp64);
int fft_cols_per_block(int n, int nxh, bool fp64);
unsigned int fft_z_blocks(const MeshGeom& g, bool fp64);
template<typename R>
hipError_t fft_forward_x(const MeshGeom& g, const FFTPlan& px, const void* tw, void* mesh, void* spec, const float* fx, hipStream_t s);
template<typename R>
hipError_t fft_inverse_x(const MeshGeom& g, const FFTPlan& px, const void* tw, const void* spec, double scale, void* phi, hipStream_t s);
template<typename R> hipError_t fft_y(const MeshGeom& g, const FFTPlan& py, const void* tw, void* spec, int sign, hipStream_t s);
template<typename R>
hipError_t fft_z_convolve(const MeshGeom& g, const FFTPlan& pz, const void* tw, void* spec, const double* G, const double* Gv,
                          int want_ev, double* ev_part, hipStream_t s);

// ---- influence function and reductions ------------------------------------
//! G = F(|k|) exp(-alpha |k|) / (k^2 |W|^2) (alpha = 0: no screening) and its virial companion Gv
hipError_t build_influence(const MeshGeom& g, int split_kind, double kappa, double rc, double k_cut, double alpha, int nq,
                           const double* qx, const double* qw, const double* w2x, const double* w2y, const double* w2z, double* G,
                           double* Gv, hipStream_t s);
hipError_t finalize(int n_ev, const double* ev_part, int n_q, const double* qpart, double inv_2V, double self_coeff, double bg_coeff,
                    double V, int want_ev, double* out, double* qsum, hipStream_t s);

// ---- short-range pair correction (EvaluatorPairPSWF over a full neighbour list) ----
//! Degree of the compact per-segment polynomial of S(s) the pair kernels evaluate: 11 in FP64;
//! 7 in FP32, whose fit error over the 64 segments (below 1e-12 in S, 1e-9 in dS/ds relative
//! to its scale) is far below FP32 resolution and which needs two 16-byte LDS loads and 14
//! FMAs per neighbour instead of three and 22.
template<typename R> constexpr unsigned int pair_compact_deg()
    {
    return sizeof(R) == 4 ? 7u : 11u;
    }
struct PairArgs
    {
    unsigned int N;
    const double* postype;
    const double* charge;
    double Lx, Ly, Lz;
    double xy, xz, yz;       // HOOMD tilt factors (0 for an orthorhombic box)
    int tilted;              // non-zero when any tilt factor is non-zero
    const unsigned int* n_neigh;
    const unsigned int* nlist;
    const uint64_t* head_list;
    const unsigned int* n_ex;
    const unsigned int* ex_list;
    unsigned int ex_pitch;
    const void* split_tab;   // [n_segs][12] reals (reference-format lookup)
    unsigned int n_segs;
    const void* compact_tab; // [n_compact][compact_deg + 1] reals of S(s), or null
    unsigned int n_compact;
    unsigned int compact_deg;
    double rc;
    double* force;           // Scalar4[N], overwritten
    double* virial;          // [6][pitch], overwritten (may be null)
    size_t virial_pitch;
    double* pq;              // scratch Scalar4[N] = (x, y, z, q), filled before the pair pass
    };
//! Pack every particle's neighbour record into a.pq: (x, y, z, q) doubles for FP64, 32-bit
//! fixed-point box fractions and a float charge (16 bytes) for FP32.
template<typename R> hipError_t pack_positions_charges(const PairArgs& a, hipStream_t s);
//! max_blocks caps the persistent grid (0 = the default of 2048 blocks).
template<typename R> hipError_t pair_forces(const PairArgs& a, int lanes_per_particle, unsigned int max_blocks, hipStream_t s);

// ---- batched solves: many independent systems per launch ------------------------------
//! Everything one system contributes to a batched launch: its geometry and plans, the
//! caller's arrays and the solver-owned tables and scratch (all device pointers).
struct BatchSys
    {
    MeshGeom g;
    FFTPlan plan[3];
    int P, D;
    unsigned int n;          // mesh-group members (group_size, or N without a group)
    unsigned int n_tiles;
    unsigned int flags;      // COMPUTE_* bits
    int want_ev;
    int ppb_x, ncol_y, ncol_z;
    unsigned int n_qblk, n_zblk;
    double inv_2V, self_coeff, bg_coeff, V;
    const unsigned int* group;
    const double* postype;
    const double* charge;
    double* force;
    const void* wcoef;
    const void* tw[3];
    const double* G;
    const double* Gv;
    void* mesh;
    void* spec;
    void* phi;
    unsigned int* tile_count;
    unsigned int* tile_start;
    unsigned int* key;
    unsigned int* slot;
    void* rec;
    SortedIdx* sidx;
    double* qpart;           // 3 per binning block: sum q, sum q^2, max |q|
    float* fx;               // fixed-point scale, its inverse, lowered flag
    FixedPointRule fxr;      // how the scan chooses the scale
    double* evpart;
    double* qsum;
    double* ev_out;
    PairArgs pair;
    };

//! One unit of batched work: (system, local block index) -- a particle chunk, a tile or an
//! FFT block, depending on the stage.
struct BatchItem
    {
    unsigned int sys;
    unsigned int local;
    };

//! Work list of one stage; `grid` blocks walk `n` items with a grid stride.
struct BatchList
    {
    const BatchItem* items;
    unsigned int n;
    unsigned int grid;
    };

constexpr int BATCH_PAIR_LANES = 8;
constexpr unsigned int BATCH_CHUNK = 256;  // particles per bin/scatter item

hipError_t batch_bin(const BatchSys* d, BatchList l, hipStream_t s);
hipError_t batch_scan(const BatchSys* d, BatchList l, hipStream_t s);
template<typename R> hipError_t batch_scatter(const BatchSys* d, BatchList l, hipStream_t s);
template<typename R> hipError_t batch_spread(const BatchSys* d, BatchList l, size_t lds, hipStream_t s);
template<typename R> hipError_t batch_gather(const BatchSys* d, BatchList l, size_t lds, hipStream_t s);
template<typename R> hipError_t batch_fft_x_r2c(const BatchSys* d, BatchList l, size_t lds, hipStream_t s);
template<typename R> hipError_t batch_fft_y(const BatchSys* d, BatchList l, int sign, size_t lds, hipStream_t s);
template<typename R> hipError_t batch_fft_z_conv(const BatchSys* d, BatchList l, size_t lds, hipStream_t s);
template<typename R> hipError_t batch_fft_x_c2r(const BatchSys* d, BatchList l, size_t lds, hipStream_t s);
//! Fused x + y passes, one workgroup per xy plane (items (system, z)), for meshes whose plane
//! fits the workgroup's LDS (fft_xy_fused); bit-identical to the separate passes.
template<typename R> hipError_t batch_fft_xy_forward(const BatchSys* d, BatchList l, size_t lds, hipStream_t s);
template<typename R> hipError_t batch_fft_xy_inverse(const BatchSys* d, BatchList l, size_t lds, hipStream_t s);
size_t fft_xy_lds(const MeshGeom& g, bool fp64);
bool fft_xy_fused(const MeshGeom& g, bool fp64);
hipError_t batch_finalize(const BatchSys* d, BatchList l, hipStream_t s);
template<typename R> hipError_t batch_pair(const BatchSys* d, BatchList l, size_t lds, bool any_tilted, hipStream_t s);
template<typename R> hipError_t batch_pack(const BatchSys* d, BatchList l, hipStream_t s);
size_t tp_lds_bytes(const MeshGeom& g, int P, bool fp64);
bool batch_order_supported(int P);
bool stencil_tile_ok(int T, int P, bool fp64, bool batched);
//! Tile edge the matrix-core FP32 spread needs at order P (17 - P) when that variant is
//! selected, else 0.
int mfma_spread_tile(int P, bool fp64);

    } // namespace kernel
    } // namespace esp
