// Molecular-dynamics support around the ESP hot path: neighbour lists, WCA, integrators.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace esp
{
namespace md
{

struct CellGeom
    {
    int nx, ny, nz;
    double Lx, Ly, Lz;
    int reach = 1;  //!< neighbour cells searched on each side: 1 (cells >= r_list) or 2 (cells >= r_list / 2)
    };

//! Optional inner list written by the same neighbour-list pass: the entries of the full list
//! with r^2 < r2 (r2 = 0: none), in rows of width nmax (HOOMD layout, head[i] = i * nmax).
struct InnerList
    {
    double r2;
    unsigned int nmax;
    unsigned int* n;
    unsigned int* list;
    uint64_t* head;
    };

//! One system of a batched MD launch (device pointers; positions, velocities and forces are
//! Scalar4 arrays).
struct MDSys
    {
    unsigned int N;
    double* postype;
    double* vel;
    double* force;
    double* virial;         // [6][pitch] or null
    size_t pitch;
    double L[3];
    const unsigned int* n_neigh;  // neighbour list (null: no WCA)
    const unsigned int* nlist;
    const uint64_t* head;
    const double* last_pos; // positions at the last list build (null: no displacement check)
    const unsigned int* n_inner;  // inner list (null: none); valid for pair cutoffs <= r_inner
    const unsigned int* inner;
    const uint64_t* head_inner;
    double r_inner;
    uint64_t seed;          // Langevin stream of this system
    };

//! (system, chunk of 256 particles)
struct MDItem
    {
    unsigned int sys;
    unsigned int chunk;
    };

//! hold (device, one byte per system, or null): systems with hold[i] != 0 are skipped
hipError_t batch_vv_first(const MDSys* d, const MDItem* items, unsigned int n_items, double dt, const unsigned char* hold, hipStream_t s);
hipError_t batch_vv_second(const MDSys* d, const MDItem* items, unsigned int n_items, double dt, double gamma, double kT, uint64_t step,
                           const unsigned char* hold, hipStream_t s);
hipError_t batch_wca(const MDSys* d, const MDItem* items, unsigned int n_items, double sigma, double epsilon, hipStream_t s);
hipError_t batch_max_disp(const MDSys* d, const MDItem* items, unsigned int n_items, unsigned long long* out, hipStream_t s);
//! dst row i = src row order[i] for n rows of `words` 4-byte words.
hipError_t gather_rows(unsigned int n, unsigned int words, const unsigned int* order, const void* src, void* dst, hipStream_t s);
hipError_t batch_predict_disp(const MDSys* d, const MDItem* items, unsigned int n_items, double dt, unsigned long long* out,
                              hipStream_t s);

hipError_t cell_count(unsigned int N, const double* pos, const CellGeom& cg, unsigned int* count, unsigned int* cell, unsigned int* slot,
                      hipStream_t s);
//! Cell member lists in ascending particle index (`unsorted`: N entries of scratch) and the
//! members' positions wrapped into the box in the same order (`spos`: N double4).
hipError_t cell_fill(unsigned int N, const double* pos, const CellGeom& cg, const unsigned int* cell, const unsigned int* slot,
                     const unsigned int* start, unsigned int* unsorted, unsigned int* members, double* spos, hipStream_t s);
//! Full (and optional inner) neighbour lists from the cell-sorted members and positions.
hipError_t build_nlist(unsigned int N, const double* spos, const CellGeom& cg, const unsigned int* cell, const unsigned int* start,
                       const unsigned int* members, double r_list, const unsigned int* n_ex, const unsigned int* ex_list,
                       unsigned int ex_pitch, unsigned int nmax, unsigned int* n_neigh, unsigned int* nlist, uint64_t* head,
                       unsigned int* overflow, const InnerList& inner, hipStream_t s);
hipError_t wca_forces(unsigned int N, const double* pos, double Lx, double Ly, double Lz, const unsigned int* n_neigh,
                      const unsigned int* nlist, const uint64_t* head, double sigma, double epsilon, double* force, double* virial,
                      size_t pitch, hipStream_t s);
hipError_t vv_first(unsigned int N, double* pos, double* vel, const double* force, double dt, double Lx, double Ly, double Lz, hipStream_t s);
hipError_t vv_second(unsigned int N, double* vel, const double* force, double dt, double gamma, double kT, uint64_t seed, uint64_t step,
                     hipStream_t s);
hipError_t max_displacement(unsigned int N, const double* pos, const double* last, double Lx, double Ly, double Lz, unsigned long long* out,
                            hipStream_t s);
hipError_t copy4(unsigned int N, const double* src, double* dst, hipStream_t s);
hipError_t energies(unsigned int N, const double* vel, const double* force, double* part, hipStream_t s);

    } // namespace md
    } // namespace esp
