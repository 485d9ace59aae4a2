// vv_internal.h -- shared declarations of the HIP library (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/vvmi.h"

namespace vv {

// ----------------------------------------------------------------- errors
void set_error(const std::string &msg);
int hip_fail(hipError_t e, const char *what);
#define VV_HIP(call)                                             \
    do {                                                         \
        hipError_t e_ = (call);                                  \
        if (e_ != hipSuccess) return ::vv::hip_fail(e_, #call);  \
    } while (0)

// ------------------------------------------------ banded elimination tables
// Thomas factors of a constant-coefficient tridiagonal system exactly as the
// reference tdma computes them (rb.f:3046-3064): fact(j) = a(j)/b'(j-1),
// b'(j) = b(j) - fact(j)*c(j-1).  Index j is 0-based.
struct TdmaTab {
    int n = 0;
    std::vector<double> fact, bp, c;
};
TdmaTab make_tdma(const std::vector<double> &a, const std::vector<double> &b,
                  const std::vector<double> &c);

// Factors of the reference cyclic solver ptdma (rb.f:3066-3116) for a
// constant-coefficient system: the r-updates use f1(j), f2(j) for
// j = 2..n-2 (1-based), then fA, fB, fC; back substitution uses the final
// a', b', c'.
struct PtdmaTab {
    int n = 0;
    std::vector<double> f1, f2, a, b, c;   // 1-based storage, size n+1
    double fA = 0, fB = 0, fC = 0;
};
PtdmaTab make_ptdma(const std::vector<double> &a, const std::vector<double> &b,
                    const std::vector<double> &c);

// Parallel-cyclic-reduction tables of one banded line operator (see
// MgLevelDev); returns the number of reduction steps.
int make_pcr(int n, const std::vector<double> &sub2, const std::vector<double> &sub1,
             const std::vector<double> &diag, const std::vector<double> &sup1,
             const std::vector<double> &sup2, std::vector<double> &tab);

// ------------------------------------------------------------- multigrid
// Device layout of one multigrid level: rows i = 0..n-1 (unknown rows of the
// line operator) contiguous, columns q = 0..ncol-1 (periodic): u[q*n + i].
// A line solve (one column) reads and writes consecutive addresses.
// lu holds the parallel-cyclic-reduction tables of the level's line
// operator: g0 | gN | alpha[S][n] | beta[S][n] | dinv[n], S = ceil(log2 n)
// (closure rows of the v operator are folded into rows 0 and n-1 first).
struct MgLevelDev {
    double *u, *f;
    double *u2;           // second iterate buffer of a smoothed level (sweeps ping-pong)
    double *dlt;          // level nlev-2: black-line changes of the last pre-sweep (n x ncol/2)
    const double *lu;
    double cphi;          // 1/hp^2 at this level
    int ncol, nstep;
    // the coarsest level is solved exactly: lu then holds one PCR table per
    // phi mode m = 0..ncol/2 (stride tabsz) followed by the twiddles
    // cos|sin(2 pi j/ncol); xm is the (ncol x n) mode scratch
    int direct;
    size_t tabsz;
    double *xm;
    const double *twid;
    // direct level with FFTs (fft_cols): the twiddles (cos, -sin)(2 pi t/n1),
    // t < n1/2, of its 2^bits = n1 point radix-2 sub-lines and, when ncol is
    // not a power of two, (cos, -sin)(2 pi t/ncol), t < ncol, of the
    // mixed-radix combination, in a buffer of their own; null elsewhere
    const double *ftw;
    // mixed precision (VV_PREC_MIXED): the level's FP32 iterate buffers,
    // right-hand side, black-line changes and PCR tables (smoothed levels);
    // on the direct level only u32, the FP32 copy of its FP64 correction
    float *u32 = nullptr, *u2_32 = nullptr, *f32 = nullptr, *dlt32 = nullptr;
    const float *lu32 = nullptr;
};
// default of vv_config.direct_cols (0): the whole plane is the direct
// level up to 8192 columns (measured faster than the V-cycle at every grid
// from 65 x 33 to 8193 x 1025), a V-cycle down to 128 columns above
inline int default_direct_cols(int np) { return np <= 8192 ? np : 128; }

// Whether a run created with configuration a can take configuration b
// (vv_reset): the same grid, geometry and solver shape -- every table of a
// run depends on these only; dt, Ra, Pr, Re, the Robert coefficient, the
// stop and the statistics switch are kernel arguments.
inline bool same_shape(const vv_config &a, const vv_config &b)
{
    auto dc = [](const vv_config &c) { return c.direct_cols > 0 ? c.direct_cols : default_direct_cols(c.np); };
    auto nu = [](int v) { return v > 0 ? v : 1; };
    return a.nr == b.nr && a.np == b.np && a.geom == b.geom && a.r0 == b.r0 && a.rmax == b.rmax &&
           (a.geom != VV_GEOM_BOX || a.width == b.width) && dc(a) == dc(b) && nu(a.nu1) == nu(b.nu1) &&
           nu(a.nu2) == nu(b.nu2) && a.precision == b.precision;
}

// a direct level of c columns is transformed by FFTs (radix 2 times the odd
// factor of c) when this holds, by dense transforms otherwise
inline bool fft_cols(int c) { return c >= 8 && c % 2 == 0 && c <= 8192; }
// the largest direct level solved with dense transforms (2 c^2 weights)
constexpr int MG_DENSE_MAXC = 4096;

constexpr int MG_MAXLEV = 16;
constexpr int MG_MAXPROB = 2;

struct MgProblemDev {
    int n;                // unknown rows
    int nlev;
    const double *sten;   // 5*n stencil of the r operator D_r (i-2..i+2)
    const double *rphi;   // n: row factor of the phi coupling (1/r^2; 1 in the box)
    const float *sten32 = nullptr, *rphi32 = nullptr;   // their FP32 copies (mixed precision)
    MgLevelDev lev[MG_MAXLEV];
};

constexpr int MG_NORM_MAXB = 1024;   // workgroups of the residual-norm kernel
constexpr int MG_HIST = 64;          // cycles of a solve whose residuals are kept

// one batch of problems solved together (v and u of one run); host-side
// description of the hierarchy and the solver state (device pointers)
struct MgBatchDev {
    int nprob;
    MgProblemDev prob[MG_MAXPROB];
    int *done;            // [nprob] converged / capped flag
    int *cycles;          // [nprob]
    unsigned long long *fnorm;  // [nprob] max |f| bits (non-negative doubles)
    double *relres;       // [nprob]
    double *hist;         // [nprob][MG_HIST] relres after each cycle
    double *npart;        // [nprob][MG_NORM_MAXB] per-workgroup residual maxima
    const double *fpart;  // [nprob][nbr] per-workgroup |f| maxima of the rhs kernel
    int nbr;
};

// Pinned host staging for the field copies between a caller's pitched
// host arrays and the device: rows are packed into (or unpacked from) it on
// the CPU, and one DMA moves it (copies from pageable memory with a pitch
// go row by row through the runtime's own staging).  Grown on demand, held
// until the run is destroyed; every copy holds the run's lock.
struct PinnedStage {
    double *p = nullptr;
    size_t n = 0;
    double *get(size_t words)
    {
        if (words <= n) return p;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
        if (hipHostMalloc(reinterpret_cast<void **>(&p), words * sizeof(double), hipHostMallocDefault) != hipSuccess) {
            p = nullptr;
            return nullptr;
        }
        n = words;
        return p;
    }
    ~PinnedStage()
    {
        if (p) (void)hipHostFree(p);
    }
};

}  // namespace vv

// ---------------------------------------------------------------- vv_run
struct vv_run {
    // every C entry point on a run holds this lock, so calls on one run from
    // several threads are serialised (the reference's main program calls
    // rhsv/poisson2 and rhsu/poisson3 in two OpenMP sections, rb.f:370-380)
    std::recursive_mutex mu;
    vv::PinnedStage stage;   // host copies (vv_put / vv_get)
    // Robert-filtered current levels waiting in Cf/thf between the explicit
    // step (vv_cvorticity / vv_temperature) and vv_update
    struct Pend { bool c = false, t = false; } pend;
    vv_config cfg;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int nr = 0, np = 0, ld = 0, ncols = 0;
    double dr = 0, dp = 0;
    size_t fbytes = 0;

    // field buffers (device, layout F[k*ld+i]); rotation by pointer swap
    double *C = nullptr, *Cp = nullptr, *Cn = nullptr, *Cf = nullptr;
    double *th = nullptr, *thp = nullptr, *thn = nullptr, *thf = nullptr;
    double *uel = nullptr, *vel = nullptr, *uelp = nullptr, *velp = nullptr;
    double *vrh = nullptr, *urh = nullptr;
    // compact-derivative scratch (the reference loans A, B, An, Bn for these)
    double *dCr = nullptr, *dCh = nullptr, *dthrC = nullptr, *dthhC = nullptr;
    double *dthrT = nullptr, *dthhT = nullptr;
    // running means (rb.f:409-417)
    double *mT2 = nullptr, *mT = nullptr, *mU2 = nullptr, *mV2 = nullptr;
    // radii of the vorticity rows r0+i*dr and temperature rows r0+(i-1/2)*dr
    // (annulus; unused in the box)
    double *rC = nullptr, *rT = nullptr;

    // elimination tables (device)
    double *tab_r1 = nullptr;   // tdma 1-4-1, n = nr-1: fact | bp | c | 1/bp
    double *tab_rT = nullptr;   // tdma with closure rows, n = nr
    double *tab_p = nullptr;    // ptdma 1-4-1, n = np+1: f1 | f2 | a | b | c | fA fB fC | 1/b

    // multigrid
    vv::MgBatchDev mg{};
    double *mg_pool = nullptr;
    double *mg_ftw = nullptr;   // FFT twiddles of the direct level (see MgLevelDev::ftw)
    float *mg_pool32 = nullptr; // FP32 hierarchy of the mixed-precision V-cycle
    double *mg_hist = nullptr;  // [2][MG_HIST] relative residual after each cycle of the last solve
    // solver state, one device block read back with a single copy:
    //   bytes 0-31 int done[2], cycles[2], unused[4];
    //   32-63 u64 unused[2], fnorm[2] (bits of non-negative doubles);
    //   64-79 double relres[2]
    void *mg_state = nullptr;
    int *mg_flags = nullptr;
    unsigned long long *mg_bits = nullptr;
    double *mg_relres = nullptr;
    int nlev_v = 0, nlev_u = 0;

    // per-step wall heat fluxes [trans1, trans2], a ring of diag_cap steps
    double *diag = nullptr;
    int diag_cap = 0;
    // per-workgroup partial reductions (no atomics): nbx = workgroups of the
    // field kernels; fpart [2][nbx] |f| maxima of the Poisson right-hand
    // sides, upart [3][nbx] CFL maxima (u, v) and clamp counts of the last
    // update, npart [2][MG_NORM_MAXB] residual maxima
    double *part = nullptr;
    double *fpart = nullptr, *upart = nullptr, *npart = nullptr;
    // kinetic energy: per-workgroup partial sums [KE_MAXB], then the total
    double *kepart = nullptr;
    int nbx = 0;
    long long steps = 0;        // completed steps since vv_start
    int last_cycles[2] = {0, 0};
    double last_relres[2] = {0, 0};
    float last_ms = 0.f;

    // per-kernel event timing (vv_profile): pending event pairs of this
    // vv_step call, resolved after its final synchronisation
    int prof_on = 0;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    struct ProfRec { int kid; hipEvent_t a, b; double bytes; };
    std::vector<ProfRec> prof_pending;
    long long prof_launch[16] = {0};
    double prof_ms[16] = {0};
    double prof_bytes[16] = {0};
};

namespace vv {
// kernel ids of the timing instrumentation (VV_K_* in include/vvmi.h)
int prof_begin(vv_run *r, hipEvent_t *a);
int prof_end(vv_run *r, int kid, hipEvent_t a, double bytes);
// explicit-step launchers (vv_explicit.hip); field_blocks = workgroups of
// the pointwise field kernels (the length of the partial-reduction arrays)
int field_blocks(const vv_run *r);
int launch_compact(vv_run *r);
int launch_df(vv_run *r, int first, int do_c, int do_t);
int launch_heat(vv_run *r, double *dst2);
int launch_rhs_guess(vv_run *r, int which, int solve);
int launch_update(vv_run *r, int nstat_kt, int from_mg);
int launch_copy_prev(vv_run *r);
// multigrid (vv_mg.hip)
int mg_build(vv_run *r);
int mg_solve(vv_run *r, int which, int *cycles, double *relres);
int mg_enqueue_deferred(vv_run *r);
int launch_mg_scatter_out(vv_run *r, int which);
int launch_compact_mask(vv_run *r, int mask);
int launch_heat_from(vv_run *r, const double *th, double *dst2);
int launch_rhs_fields(vv_run *r, int which, int solve);
// kinetic energy of uel, vel into kepart[KE_MAXB]
constexpr int KE_MAXB = 1024;
int launch_energy(vv_run *r);
}  // namespace vv
