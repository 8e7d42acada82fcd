// qjx_rk.h -- launcher of the Runge-Kutta trajectory kernel (qjx_rk.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qjx_internal.h"

namespace qjx {

// Band coefficient arrays of the RK right-hand side, each P E doubles in
// lane-major order: element i = l E + j (lane l of a trajectory's P, slot j) sits at
// (a E + j) P + l.
enum RkBand {
    kRkDre = 0,   // Re H_part(i, i)
    kRkDim,       // Im H_part(i, i)
    kRkO1re,      // Re H_part(i, i+1) = Re H_part(i+1, i)
    kRkO1im,      // Im H_part(i, i+1)
    kRkO2im,      // Im H_part(i, i+2) (the real part is an exact zero)
    kRkDrv,       // H_drv(i, i)
    kRkLd,        // L(i, i)          (jumps only)
    kRkLm,        // L(i, i-1)
    kRkLp,        // L(i, i+1)
    kRkBands
};
constexpr int kRkStageBands = kRkLd;   // arrays staged in LDS (the stencil's)

// Lanes per trajectory: 32 (two trajectories per wavefront) for 64 < d <= 128, where
// one wavefront of two entries per lane would spend as much on its cross-lane halo
// and norm as on the stencil; 64 otherwise.
inline int rk_group(int d) { return d > 64 && d <= 128 ? 32 : 64; }

// Entries per lane: the smallest power of two with P E >= d (one entry per lane up
// to d = 64, four from d = 65 with 32 lanes).
inline int rk_lanes(int d)
{
    const int P = rk_group(d);
    int E = 1;
    while (P * E < d) E *= 2;
    return E;
}

struct RkArgs {
    const double *bands;      // [kRkBands][E][P]
    double *tab;              // [3 nsteps]: drives of each step's full integration (scratch)
    void *slot0, *slot1;      // FP64 state columns (L doubles: Re at [0, dp), Im at [dp, 2 dp));
                              // dp = P E, the padding beyond d zero
    TrajState *st;
    JumpLog log;
    int *err;                 // [4]: code, trajectory, bits of the offending norm
    unsigned long long *counters;  // [0] right-hand-side evaluations, [1] jumps
    uint64_t stride;          // MCG59 leap-frog multiplier
    long long pos0;           // steps done before this launch (period = pos / rk_ns)
    double T, rk_step, prm_E, ampl, freq, phase;
    int M, d, dp, L;
    int nsteps, rk_ns, drv_type;
    int jump, is_obs;
};

// Advance every trajectory by a.nsteps RK steps (P lanes per trajectory);
// a.tab must hold 3 a.nsteps doubles.  Returns hipErrorInvalidValue if d exceeds
// the kernel's largest lane width.
hipError_t launch_rk(const RkArgs &a, hipStream_t s);

}  // namespace qjx
