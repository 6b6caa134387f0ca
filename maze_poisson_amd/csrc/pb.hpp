// pb.hpp -- Poisson-Boltzmann boundary built from ion radii: the dielectric and
// screening maps (rebuilt every step from the ion positions) and the
// dielectric-boundary, ionic-boundary and non-polar forces.
#pragma once

#include "common.hpp"

namespace maze {

constexpr int kPbTile = 8;     // edge of the grid tiles and ion bins of the map kernel
constexpr int kPbCap = 2048;   // candidate ions one map tile can hold
constexpr int kPbMaxSeg = 512; // bin-row segments one map tile can scan
constexpr int kPbChunk = 256;  // candidates staged in LDS at a time (= threads)

struct PBDev {
    int np = 0, n = 0;
    int active = 0;       // radii given: maps and forces available
    int nonpolar = 0;
    double h = 0, w = 0, eps_s = 1, eps_int = 1, kbar2 = 0, gamma_np = 0, beta_np = 0;
    double* radii = nullptr;  // (np) Bohr
    double* rp2 = nullptr;    // (r + w)^2
    double* rm2 = nullptr;    // (r - w)^2
    int* rmap = nullptr;      // maps box half-width floor((r + w) / h) + 1
    int* rfrc = nullptr;      // forces box half-width floor((r + w) / h) + 2
    int rmax = 0;             // max of rmap
    int nb = 0, key_bits = 1; // bins per axis, bits of a bin key
    unsigned int* keys = nullptr;
    unsigned int* keys_sorted = nullptr;
    int* ids = nullptr;
    int* ids_sorted = nullptr;
    void* sort_tmp = nullptr;
    size_t sort_bytes = 0;
    double* fdb = nullptr;    // (np, 3) dielectric-boundary force
    double* fib = nullptr;    // (np, 3) ionic-boundary force
    double* fnp = nullptr;    // (np, 3) non-polar force
    int* flag = nullptr;      // device: 1 when a map tile overflowed kPbCap / kPbMaxSeg
};

// Bins the ions by their cell and sorts them (stable: ion order within a bin).
void launch_pb_bin(hipStream_t s, PBDev& b, const double* pos);
// eps_x/y/z and k2 of the current ion positions (needs launch_pb_bin first).
void launch_pb_maps(hipStream_t s, const PBDev& b, const double* pos, double* ex, double* ey,
                    double* ez, double* k2);
// fdb, fib, fnp from the field and the maps.
void launch_pb_forces(hipStream_t s, const PBDev& b, const double* pos, const double* phi,
                      const double* ex, const double* ey, const double* ez, const double* k2);
size_t pb_sort_bytes(int np, int key_bits);

}  // namespace maze
