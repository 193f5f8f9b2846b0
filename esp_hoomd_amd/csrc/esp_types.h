// POD descriptors shared by the host orchestration and the HIP kernels.
#pragma once

#include <stdint.h>

namespace esp
{

constexpr int MAX_FFT_STAGES = 16;
constexpr int MAX_ORDER = 16;      // largest stencil order P with a kernel instantiation
constexpr int MAX_WINDOW_DEG = 15; // largest window polynomial degree

//! Mixed-radix (4, 2, 3, 5) Stockham plan for one 1-D length.
struct FFTPlan
    {
    int n;
    int nstages;
    int radix[MAX_FFT_STAGES];
    };

//! Geometry of one periodic system and its mesh.
struct MeshGeom
    {
    int nx, ny, nz;     // mesh points per axis
    int nxh;            // nx / 2 + 1 (half spectrum along x)
    int sxp;            // row pitch of the half spectrum and the influence tables: nxh rounded
                        // up to 16 elements, so 16-column row segments are 128-byte aligned
    int T;              // spreading tile edge (cells)
    int tx, ty, tz;     // tiles per axis
    int B;              // LDS box edge T + P - 1
    double Lx, Ly, Lz;  // box lengths (box centred on the origin)
    double hx, hy, hz;  // L / n per axis (the mesh axes follow the lattice vectors)
    double xy, xz, yz;  // HOOMD tilt factors: lattice vectors a1 = (Lx, 0, 0),
                        // a2 = (xy Ly, Ly, 0), a3 = (xz Lz, yz Lz, Lz)
    int tilted;         // non-zero when any tilt factor is non-zero
    };

//! Particle record in tile-sorted order (one per charged group member).
template<typename R> struct SortedRec
    {
    R tx, ty, tz, q;  // stencil fractional offsets t in [0,1) and charge
    };

struct SortedIdx
    {
    uint32_t local;   // origin within the tile: ox | oy << 8 | oz << 16
    uint32_t idx;     // particle index into postype / force
    };

//! How the tile scan chooses the mesh's fixed-point scale each step (fixed_point_scale in
//! kernels_mesh.hip): 2^23 / 2^ceil(log2 max|q|) for FP32 taps into int32 LDS cells, 2^50 /
//! 2^ceil(log2 max|q|) for FP64 taps into int64 cells, halved while the most populated tile
//! could overflow the accumulators.
struct FixedPointRule
    {
    float w3max;  // largest |w_x w_y w_z| of the window
    float cover;  // tile boxes that overlap one mesh cell: (1 + ceil((P - 1) / T))^3
    int wide;     // 0: FP32 taps into int32 LDS cells; 1: FP64 taps into int64 cells
    };
//! Tile boxes (edge T + P - 1, one per tile of edge T) that contain a given mesh cell.
inline float tile_box_cover(int T, int P)
    {
    const int per_axis = 1 + (P - 1 + T - 1) / T;
    return float(per_axis * per_axis * per_axis);
    }

    } // namespace esp
