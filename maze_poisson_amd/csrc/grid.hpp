// grid.hpp -- device state and launchers for the grid solvers.
#pragma once

#include "common.hpp"

namespace maze {

constexpr int kMaxPasses = 16;  // refinement passes recorded per solve

// Scalars of one Krylov solve, resident in device memory so that the
// per-iteration kernels never wait for the host.  Written only by the last
// work-group of each reduction pass.
struct CGState {
    double rdv;      // <r, v> of the current direction (v = -r/6)
    double pAp;      // <p, A p>
    double alpha, beta;
    double rr, rmax; // <r, r> and max|r| after the last update
    double tol;      // absolute stop on max|r|
    double rtol;     // relative stop on max|r| / max|r0| (0: absolute mode)
    double r0;       // max|r| of the initial residual
    double resid;    // spare scalar (true-residual passes)
    int iter;        // 0-based index of the iteration in flight
    int limit;       // iteration cap
    int done;        // 1 once stopped (converged, capped or broken down)
    int res;         // result code (see cg_result)
    int first;       // 1 until the first direction has been formed
    int count_mode;  // 1: result is the number of iterations (coarse CG)
    int pcg;         // 1: beta comes from the preconditioner pass, not from <r, r>
    int plain;       // 1: plain CG (p = -r), else the implicit-Jacobi recurrence (p = r/6)
    // Outer FP64 refinement loop of the mixed-precision solve, advanced by the
    // refinement pass so that a whole predicted schedule of passes can be
    // queued without host round trips (passes after convergence are no-ops).
    int outer_done;  // 1 once max|sigma - A y| <= otol
    int ysel;        // which of the two y buffers holds the current iterate
    int passes;      // refinement passes applied
    int inner_total; // inner iterations summed over the applied passes
    int pass_iters[kMaxPasses];  // per pass: iterations if converged, else -(iterations)
    double otol;     // absolute stop of the outer loop
    // <r, z> of the multigrid-PCG inner iterations by iteration parity: the
    // direction pass of iteration k writes slot k & 1 and reads the other.
    double rdv2[2];
};

// Partials of the multigrid-PCG inner iteration's reductions, folded by the
// kernel that consumes them instead of by a last-arriving work-group
// (deferred finalisation: no arrival counter, no tail on the producer):
// <r, z> of the finest up kernel (pz), <p, A p> of the direction pass (pp),
// <r, r> and max|r| of the update (pus, pum).
struct DeferRed {
    double* pz;
    double* pp;
    double* pus;
    double* pum;
};

// What the host reads after a solve: the solver state and the charge sum of
// the last spread, written by launch_publish_state into host-coherent memory
// with `seq` last, so that the host can wait by polling seq instead of
// synchronising the stream.
struct StatePub {
    CGState st;
    double q_sum;
    unsigned int seq;
};
void launch_publish_state(hipStream_t s, const CGState* st, const double* q_sum, StatePub* pub,
                          unsigned int seq);

// Face-centred dielectric maps and screening term of the PB operator; the
// Poisson operator passes nullptrs.  The FP32 form serves the FP32 inner
// solves of the PB multigrid path.
template <class T>
struct PBCoefT {
    const T* ex;
    const T* ey;
    const T* ez;
    const T* k2;
};
using PBCoef = PBCoefT<double>;
using PBCoef32 = PBCoefT<float>;

struct Reduce {
    double* psum;
    double* pmax;
    unsigned int* counter;
};

// ---- launchers (grid_kernels.hip) ----------------------------------------
void launch_laplace(hipStream_t s, const GridDims& d, const PBCoef& pb, const double* u, double* out);
void launch_pred_sigma(hipStream_t s, const GridDims& d, const PBCoef& pb, const double* phi,
                       const double* phip, const double* q, double coef, double* phistar,
                       double* sigma);
template <class Real>
void launch_cg_init(hipStream_t s, const GridDims& d, const PBCoef& pb, const Real* x,
                    const Real* b, Real* r, CGState* st, Reduce red, Real* zero_x = nullptr,
                    int plain = 0);
template <class Real>
void launch_cg_dir(hipStream_t s, const GridDims& d, const PBCoef& pb, const Real* p_old,
                   const Real* r, Real* p_new, CGState* st, Reduce red, int first);
// zero_x: x is taken as 0 and only written (first update from a zero start).
template <class Real>
void launch_cg_update(hipStream_t s, const GridDims& d, const PBCoef& pb, const Real* p, Real* x,
                      Real* r, CGState* st, Reduce red, int zero_x = 0);
// <r, v> of the explicit-preconditioner recurrence (library/src/plcg.c:127,
// :146-148): first: r_dot_v = <r, v>; else beta = <r, v> / r_dot_v and
// r_dot_v = <r, v>.
void launch_dot_rv(hipStream_t s, const GridDims& d, const double* r, const double* v, CGState* st,
                   Reduce red, int first);
// gated: the solve is started only while the outer refinement loop runs.
void launch_cg_reset(hipStream_t s, CGState* st, double tol, double rtol, int limit, int count_mode,
                     int pcg = 0, int gated = 0, int plain = 0);

// Mixed precision, FP64 true residual of the refinement loop, b = sigma - A y.
// First pass (add_d == 0): y = y0, starts the loop (ysel = 0).
// Later passes: y_new = y + d with y/y_new the two buffers y0/y1 selected by
// st->ysel (flipped); skipped once the loop is done.  Each pass stops the
// loop when max|b| <= otol, else starts the next inner solve A x = b from
// x = 0 (r32 = -b in FP32, relative stop rtol, iteration cap limit, pcg as
// in CGState).
struct RefineArgs {
    double otol, rtol;
    int limit, pcg;
};
// defer: the inner solve before this pass ended with a deferred update whose
// stop test the refinement pass applies first (nullptr: none pending).
void launch_refine(hipStream_t s, const GridDims& d, const PBCoef& pb, double* y0, double* y1,
                   const float* dcorr, const double* sigma, float* r32, RefineArgs a, CGState* st,
                   Reduce red, int add_d, const DeferRed* defer = nullptr);
// First pass from the extrapolated iterate y_out = 2 y - y_prev (then y_out is
// the loop's y0).
void launch_refine_extrap(hipStream_t s, const GridDims& d, const PBCoef& pb, const double* y,
                          const double* y_prev, double* y_out, const double* sigma, float* r32,
                          RefineArgs a, CGState* st, Reduce red);

void launch_axpy(hipStream_t s, long n, double alpha, const double* v, double* u);
// phi -= (st->ysel ? yb : ya) if st->outer_done, else nothing (the same
// values as launch_axpy with alpha -1 on the selected buffer).
void launch_phi_update_gated(hipStream_t s, long n, const CGState* st, const double* ya,
                             const double* yb, double* phi);
void launch_scale_copy(hipStream_t s, long n, double c, const double* in, double* out);
// Block-Jacobi preconditioner v = A_B^-1 r on every 10^3 block (n % 10 == 0);
// tab: the blockjacobi_tables() values (1100 doubles) in device memory.
constexpr int kBlockJacobiTab = 1100;
void blockjacobi_tables(double* tab);
void launch_blockjacobi(hipStream_t s, int n, const double* in, double* out, const double* tab);
void launch_fill(hipStream_t s, long n, double v, double* out);
void launch_fill_f(hipStream_t s, long n, float v, float* out);
void launch_copy(hipStream_t s, long n, const double* in, double* out);
void launch_half_sum(hipStream_t s, long n, const double* phi, double* out, Reduce red);
void launch_dot_scaled(hipStream_t s, long n, double scale, const double* a, const double* b,
                       double* out, Reduce red);
void launch_sum_scaled(hipStream_t s, long n, double scale, const double* v, double* out,
                       Reduce red);

// Reference two-grid multigrid pieces.
void launch_rbgs(hipStream_t s, const GridDims& d, const double* in, double* out, int iters);
void launch_mg_residual(hipStream_t s, const GridDims& d, const PBCoef& pb, const double* in,
                        const double* out, double* r);
// Reference Poisson-Boltzmann multigrid pieces (library/src/multigrid_pb.c).
void launch_rbgs_pb(hipStream_t s, const GridDims& d, const PBCoef& pb, const double* in,
                    double* out, int iters);
void launch_restrict_pb_coef(hipStream_t s, int n, const PBCoef& fine, double* ex, double* ey,
                             double* ez, double* k2);
void launch_restrict27(hipStream_t s, int n, const double* in, double* out);
void launch_prolong_add(hipStream_t s, int nc, const double* in, double* out);
void launch_true_residual_max(hipStream_t s, const GridDims& d, const PBCoef& pb, const double* y,
                              const double* sigma, CGState* st, Reduce red);

// FP32 multigrid preconditioner (mg_kernels.hip).
void launch_mg_pre(hipStream_t s, const GridDims& d, const float* b, float* dout, double c,
                   CGState* st);
void launch_mg_resid(hipStream_t s, const GridDims& d, const float* b, const float* din, float* r,
                     CGState* st);
void launch_mg_post(hipStream_t s, const GridDims& d, const float* b, const float* din,
                    float* dout, double c, CGState* st, int dot, Reduce red);
void launch_mg_restrict(hipStream_t s, int n, const float* r, float* bc, CGState* st);
void launch_mg_prolong_add(hipStream_t s, int nc, const float* ec, float* d, CGState* st);
// Exact solve on the coarsest level: one work-group with the grid in LDS for
// c <= kMgCoarseLds, else six
// axis passes over the work arrays w0, w1 (c^3 doubles each) for
// c <= kMgCoarseMax.
constexpr int kMgCoarseLds = 16;
constexpr int kMgCoarseMax = 128;
void launch_mg_coarse(hipStream_t s, int c, const float* b, float* e, const double* Qf,
                      const double* ilam, CGState* st);
void launch_mg_coarse(hipStream_t s, int c, const double* b, double* e, const double* Qf,
                      const double* ilam, CGState* st);
// The down kernel of the level above a 16^3 coarsest grid with the coarse
// solve in its last work-group (count: a zeroed device counter the kernel
// leaves zeroed; e: the coarse correction).  Returns false, launching nothing,
// where that form is not built (the caller then launches the two kernels).
struct CoarseTail {
    unsigned int* count;
    const double* Qf;
    const double* ilam;
    float* e;
};
bool mg_down_coarse_ok(int n, int tile, int c);
bool launch_mg_down_coarse(hipStream_t s, int n, int tile, const float* b, float* d2, float* bc,
                           double c1, double c2, CGState* st, int c, const CoarseTail& tail);
void launch_mg_coarse_big(hipStream_t s, int c, const float* b, float* e, double* w0, double* w1,
                          const double* Qf, const double* ilam, CGState* st);
// The multigrid-PCG direction and update with deferred finalisation (Poisson
// operator).  pcg_dir: beta from the finest up kernel's partials (zcnt of
// them; 0 for the first direction), writes <r, z> to rdv2[par] and its
// <p, A p> partials; after the first iteration it first applies the previous
// update's stop test.  update: alpha = rdv2[par] / <p, A p> from those
// partials, x += alpha p, r += alpha A p, partials of <r, r> and max|r| for
// the next consumer's stop test (the next direction pass, the refinement
// pass or launch_df_finalize).
void launch_pcg_dir_df(hipStream_t s, const GridDims& d, const float* p_old, const float* v,
                       float* p_new, CGState* st, const DeferRed& dr, int zcnt, int first, int par);
void launch_cg_update_df(hipStream_t s, const GridDims& d, const float* p, float* x, float* r,
                         CGState* st, const DeferRed& dr, int zero_x, int par);
// The same in FP64 (the FP64 multigrid-PCG).
void launch_pcg_dir_df(hipStream_t s, const GridDims& d, const double* p_old, const double* v,
                       double* p_new, CGState* st, const DeferRed& dr, int zcnt, int first, int par);
void launch_cg_update_df(hipStream_t s, const GridDims& d, const double* p, double* x, double* r,
                         CGState* st, const DeferRed& dr, int zero_x, int par);
// Applies the deferred stop test of the last update (one work-group): for a
// host read of the state in the middle of an inner solve.
void launch_df_finalize(hipStream_t s, CGState* st, const DeferRed& dr, int cnt);
void launch_pcg_dir(hipStream_t s, const GridDims& d, const float* p_old, const float* v,
                    float* p_new, CGState* st, Reduce red, int first, const PBCoef& pb = PBCoef{});
// FP32 multigrid for the linear PB operator (mg_kernels.hip), with FP32
// coefficients on every level.
void launch_pcg_dir(hipStream_t s, const GridDims& d, const float* p_old, const float* v,
                    float* p_new, CGState* st, Reduce red, int first, const PBCoef32& pb);
void launch_cg_update32(hipStream_t s, const GridDims& d, const PBCoef32& pb, const float* p,
                        float* x, float* r, CGState* st, Reduce red, int zero_x);
void launch_pb_smooth(hipStream_t s, const GridDims& d, const PBCoef32& pb, const float* g,
                      const float* b, const float* din, float* dout, double c, CGState* st,
                      int dot, Reduce red);
void launch_pb_pre(hipStream_t s, const GridDims& d, const PBCoef32& pb, const float* g,
                   const float* b, float* dout, double c1, double c2, CGState* st);
void launch_pb_resid(hipStream_t s, const GridDims& d, const PBCoef32& pb, const float* b,
                     const float* din, float* r, CGState* st);
void launch_pb_to32(hipStream_t s, long count, const PBCoef& in, float* ex, float* ey, float* ez,
                    float* k2);
void launch_pb_idg(hipStream_t s, int n, const PBCoef32& pb, float* g);
void launch_pb_coarsen(hipStream_t s, int nf, const PBCoef32& f, float* cx, float* cy, float* cz,
                       float* ck2);
void launch_pb_coarse(hipStream_t s, int c, const PBCoef32& pb, const float* g, const float* b,
                      float* x, int iters, double lo, double hi, CGState* st);
// The level above the coarsest (n <= 20, even) and the coarsest (n / 2) of the
// PB V-cycle in one work-group: reads that level's b, writes its d; the same
// values as the separate passes (pre-smoothing with c1 then c2, residual,
// restriction, `iters` coarsest iterations on [lo, hi], prolongation, post-
// smoothing with c2 then c1).
constexpr int kPbTailMaxN = 20;
void launch_pb_tail(hipStream_t s, int n, const PBCoef32& pf, const float* gf, const float* b,
                    float* d, const PBCoef32& pc, const float* gc, double c1, double c2, int iters,
                    double lo, double hi, CGState* st);
// Fused level kernels of the V-cycle on an n^3 level (n even):
// down: d2 = two pre-smoothing sweeps from zero, b_coarse = restricted residual;
// up:   d = two post-smoothing sweeps of d2 + prolong(e); dot: also <b, d> and
//       the PCG beta update (finest level).
// `tile` selects the work-group tile shape (MgTile).
enum MgTile {
    kMgT8x16x16_1024 = 0,
    kMgT8x16x16_512,
    kMgT8x8x16_512,
    kMgT4x8x8_512,
    kMgT4x8x8_256,
    kMgMarch16,  // plane march, 16x16 columns, 16 planes per work-group (n % 16 == 0)
    kMgMarch8,   // plane march, 16x16 columns, 8 planes per work-group (n % 16 == 0)
    kMgTiles
};
// Or-ed into a tile id below kMgMarch16: the same tile shape with the scalar-load
// kernels that serve odd and small levels, in place of the column-mapped
// vector-load kernels (the two compute every value alike).
constexpr int kMgScalarLoads = 64;
int mg_default_tile(int n);     // down kernel
int mg_default_tile_up(int n);  // up kernel
// Whether `tile` can run a level of n^3 points (the launchers fall back to the
// default tile otherwise).
bool mg_tile_ok(int tile, int n);
const char* mg_tile_name(int tile);
// Smoother weights c1, c2 (damping / 6): pre-smoothing sweeps c1 then c2,
// post-smoothing c2 then c1 (polynomials in A, so the cycle stays symmetric).
void launch_mg_down(hipStream_t s, int n, int tile, const float* b, float* d2, float* bc,
                    double c1, double c2, CGState* st);
// dot: the finest level also writes its <b, d> partials to red.psum (one per
// work-group, folded by the PCG direction pass; mg_up_blocks counts them).
void launch_mg_up(hipStream_t s, int n, int tile, const float* b, const float* d2, const float* e,
                  float* d, double c1, double c2, CGState* st, int dot, Reduce red);
int mg_up_blocks(int n, int tile);
// The level kernels in FP64 (a fixed tile shape per level size); the finest up
// kernel's partials number mg_up_blocks64(n).
void launch_mg_down(hipStream_t s, int n, const double* b, double* d2, double* bc, double c1,
                    double c2, CGState* st);
void launch_mg_up(hipStream_t s, int n, const double* b, const double* d2, const double* e,
                  double* d, double c1, double c2, CGState* st, int dot, Reduce red);
int mg_up_blocks64(int n);
// Rest of the V-cycle from a small level (sizes n[0..nlev), finest first) down
// to the exact solve and back, in one work-group with the data in LDS:
// reads that level's right-hand side b, writes its correction d.
bool mg_tail_fits(int nlev, const int* n);
void launch_mg_tail(hipStream_t s, int nlev, const int* n, const float* b, float* d, double c1,
                    double c2, const double* Qf, const double* ilam, CGState* st);

}  // namespace maze
