// particles.hpp -- device particle state and launchers.
#pragma once

#include "common.hpp"
#include "grid.hpp"

namespace maze {

struct PDev {
    int np, n_typ, n;
    int cas, nn;      // charge-assignment scheme (0 CIC, 1/2 quadratic/cubic spline), nodes per ion
    double L, h, r_cut;
    int* types;
    double* pos;      // (np, 3) Bohr
    double* vel;      // (np, 3)
    double* mass;     // (np)
    double* charges;  // (np)
    double* fe;       // electrostatic force (np, 3)
    double* fn;       // Tosi-Fumi force (np, 3)
    double* ft;       // total force (np, 3)
    int* base;        // CIC base cell (np, 3), wrapped into [0, n)
    // Charge spread: per grid node the head of a list of the (ion, node)
    // contributions it receives, tagged with the spread's epoch (high 32 bits;
    // entries of an older epoch count as empty, so the array is never cleared);
    // per contribution its grid node and the next entry of its node's list.
    unsigned long long* head;  // (n^3)
    int* node;        // (np, nn)
    int* link;        // (np, nn)
    double* table;    // Tosi-Fumi constants per type pair, 7 each
    double* tpart;    // Tosi-Fumi partials per 64-ion block pair (tf_part_doubles)
    double* eion;     // Tosi-Fumi energy per ion (each pair's energy on both ions)
    int* counts;      // pairs within the cutoff per ion
    double* contrib;  // per-(ion, node) assignment contribution, (np, nn)
};

// Device-resident scalars of the particle side.
struct PScalars {
    double q_sum;      // sum of all CIC contributions of the last spread
    double pe_tf;      // Tosi-Fumi potential energy
    double kinetic;    // 1/2 sum m v^2
    double scale;      // velocity-rescale factor of the Verlet thermostat
    double mom[3];
    double typesum[3 * 16];
    int pb_flag;       // set by the PB map kernel when a tile overflows
    unsigned int epoch;  // charge spreads completed (tags the node lists)
};

void launch_cic_base(hipStream_t s, const PDev& p);
// The charge spread onto q (q must hold zeros except at the nodes of the
// previous spread, which zero_old clears; without zero_old q must be all zero).
void launch_assign(hipStream_t s, const PDev& p, const PScalars* ps, double* q, int zero_old);
// Sum of all contributions into ps->q_sum; advances ps->epoch.
void launch_contrib_sum(hipStream_t s, const PDev& p, PScalars* ps, Reduce red);
void launch_force_fd(hipStream_t s, const PDev& p, const double* phi);
// The MD step's path: the field force, F_tot = F_elec + F_noel and the
// velocity half-kick v += cdt F_tot / m in one launch (no PB boundary forces).
// gate: if set, the launch does nothing unless *gate is nonzero on the device
// when it runs (work queued before the host knows whether it is due).
void launch_force_fd_kick(hipStream_t s, const PDev& p, const double* phi, double cdt,
                          const int* gate = nullptr);
long tf_part_doubles(int np);
void launch_tf(hipStream_t s, const PDev& p, PScalars* ps, Reduce red);
// The same in pieces: the 64-ion block pairs (I, J >= I) of rows I in [i_lo, i_hi),
// then the per-ion fold and the energy (after every row has run).
int tf_blocks(int np);
void launch_tf_rows(hipStream_t s, const PDev& p, int i_lo, int i_hi);
void launch_tf_finish(hipStream_t s, const PDev& p, PScalars* ps, Reduce red);
// F_tot = F_elec + F_noel (+ the PB boundary forces fdb, fib, fnp when given).
void launch_force_total(hipStream_t s, const PDev& p, const double* fdb = nullptr,
                        const double* fib = nullptr, const double* fnp = nullptr);
void launch_kinetic(hipStream_t s, const PDev& p, PScalars* ps, Reduce red, double T_target,
                    int want_scale);
// with_base: also the base cells of the new positions (launch_cic_base's work).
void launch_verlet_part1(hipStream_t s, const PDev& p, double dt, const PScalars* ps, int rescale,
                         bool with_base = false);
void launch_kick(hipStream_t s, const PDev& p, double cdt);
void launch_drift(hipStream_t s, const PDev& p, double c2dt);
void launch_o_block(hipStream_t s, const PDev& p, uint64_t seed, uint64_t call, double c1s,
                    double var2);
void launch_momentum(hipStream_t s, const PDev& p, PScalars* ps);
void launch_rescale_velocities(hipStream_t s, const PDev& p, PScalars* ps);

}  // namespace maze
