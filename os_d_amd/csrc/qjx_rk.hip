// qjx_rk.hip -- Runge-Kutta quantum-jump trajectories of the dimer (os_d prop_id 1).
//
// Reference (source/cpp/QJX/QJX/): rk_period (core.cpp:439-450), rk_step_dimer
// (core.cpp:11886-11948), rk_int_dimer (core.cpp:11592-11619), rk_right_part_dimer
// (core.cpp:11033-11115), arg_upd / rk_final / rk_recovery / restore_from_prev
// (rk_proc.cpp:141-345).
//
// The reference evaluates k = -i (H_part + E(t) H_drv) phi with two dense zgemv
// per RK stage.  H_part = H - i/2 gamma' L^dag L is pentadiagonal and H_drv is
// diagonal, so every other term is an exact zero: the stage is a five-point
// stencil, bandwidth- and latency-bound, not a GEMM.  P lanes own one trajectory
// for a whole launch (P = 64, or 32 for 64 < d <= 128: two trajectories per
// wavefront): lane l of the group holds entries [l E, l E + E) of phi, of the stage
// argument and of the RK accumulator in registers; the stencil's neighbours across
// lanes come from four cross-lane shuffles per stage (wavefront-wide: at the edge
// between two trajectories the shifted-in values meet band coefficients that are
// exact zeros, those of columns outside 0..d-1); the band coefficients (shared by
// all trajectories) sit in LDS, lane-major so a wavefront's reads are
// conflict-free; ||phi||^2 is a lane sum plus a butterfly over the group.
// Nothing touches HBM between the first load and the final store of the state
// except jump records.
//
// Arithmetic order (bit-identical to the CPU restatement): one fma per
// multiply-add, terms in ascending column order, RK4 update phi + acc (step/6),
// ||x||^2 summed per lane then by the butterfly over lane distances P/2..1.
// Nothing else is contracted.
#include <hip/hip_runtime.h>

#include "qjx_device.h"
#include "qjx_rk.h"

#pragma clang fp contract(off)

namespace qjx {
namespace {

constexpr int kWaves = 4;   // trajectories (wavefronts) per workgroup
constexpr int kRkMaxCrossings = 1 << 16;   // per RK step; more is an error

// Taylor coefficients (-1)^n / (2n+1)! for n = 11..1, then (-1)^n / (2n)! for n = 12..1.
__constant__ double kRkPoly[23] = {
    -0x1.761b41316381ap-75, 0x1.71b8ef6dcf572p-66, -0x1.2f49b46814157p-57, 0x1.952c77030ad4ap-49,
    -0x1.ae7f3e733b81fp-41, 0x1.6124613a86d09p-33, -0x1.ae64567f544e4p-26, 0x1.71de3a556c734p-19,
    -0x1.a01a01a01a01ap-13, 0x1.1111111111111p-7, -0x1.5555555555555p-3,
    0x1.f2cf01972f578p-80, -0x1.0ce396db7f853p-70, 0x1.e542ba4020225p-62, -0x1.6827863b97d97p-53,
    0x1.ae7f3e733b81fp-45, -0x1.93974a8c07c9dp-37, 0x1.1eed8eff8d898p-29, -0x1.27e4fb7789f5cp-22,
    0x1.a01a01a01a01ap-16, -0x1.6c16c16c16c17p-10, 0x1.5555555555555p-5, -0x1.0000000000000p-1};

// sin x: quadrant k = nearest integer to x 2/pi, r = x - k pi/2 with pi/2 split
// into 33 + 33 + 53 significant bits (Cody & Waite 1980), then the Taylor series of
// sin r or cos r on |r| <= pi/4 by Horner.  Same operations as the CPU restatement.
// The coefficients are scalar loads through an offset the compiler cannot see
// through, so they are not kept in vector registers across the step loop.
__device__ inline double rk_sin(double x)
{
    int b = 0;
    asm volatile("" : "+s"(b));
    const double *P = kRkPoly + b;
    const double k = __builtin_rint(x * 0x1.45f306dc9c883p-1);
    double r = x - k * 0x1.921fb54400000p+0;
    r = r - k * 0x1.0b4611a600000p-34;
    r = r - k * 0x1.3198a2e037073p-69;
    const double z = r * r;
    double ps = P[0];
#pragma unroll
    for (int n = 1; n < 11; n++) ps = __builtin_fma(ps, z, P[n]);
    double pc = P[11];
#pragma unroll
    for (int n = 12; n < 23; n++) pc = __builtin_fma(pc, z, P[n]);
    const double s = __builtin_fma(r * z, ps, r);
    const double c = __builtin_fma(z, pc, 1.0);
    const int q = (int)((long long)k & 3);
    return q == 0 ? s : q == 1 ? c : q == 2 ? -s : -c;
}

// E(t) (core.cpp:11061-11088): type 1 is E + A sin(omega t + phase), any other the
// square drive E + A on the first half period, E - A on the second.
__device__ inline double rk_drive(const RkArgs &a, double t)
{
    if (a.drv_type == 1) return a.prm_E + a.ampl * rk_sin(a.freq * t + a.phase);
    const double mt = fmod(t, a.T);
    return mt < a.T * 0.5 ? a.prm_E + a.ampl : a.prm_E - a.ampl;
}

// Neighbour exchange by wavefront shifts (DPP wave_shr:1 / wave_shl:1, one
// v_mov_dpp per dword): shfl_up1 gives lane l the value of lane l-1 (lane 0: 0),
// shfl_down1 the value of lane l+1 (lane 63: 0).
__device__ inline double shfl_up1(double v, int)
{
    const unsigned long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffULL), 0x138, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x138, 0xf, 0xf, true);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

__device__ inline double shfl_down1(double v, int)
{
    const unsigned long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffULL), 0x130, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x130, 0xf, 0xf, true);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// ||x||^2 of the group's trajectory: lane sums, then the butterfly over lane
// distances P/2..1 (every lane of the group gets the same bits)
template <int E, int P>
__device__ inline double group_norm(const double (&xr)[E], const double (&xi)[E])
{
    double p = 0.0;
#pragma unroll
    for (int j = 0; j < E; j++) {
        p = __builtin_fma(xr[j], xr[j], p);
        p = __builtin_fma(xi[j], xi[j], p);
    }
    if constexpr (P == 64) {
        return wave_sum(p);
    } else {
        for (int off = P / 2; off > 0; off >>= 1) p += __shfl_xor(p, off, 64);
        return p;
    }
}

// Coefficients of the stencil this lane needs from its left neighbour lane:
// H_part(lE-1, lE), and the imaginary H_part(lE-2, lE), H_part(lE-1, lE+1).
struct Halo {
    double o1re, o1im, o2a, o2b;
};

enum { kStageFirst = 1, kStageMid = 2, kStageLast = 4 };

// One RK stage over the lane's entries: k = f(x) with drive Ed, then
//   kStageFirst: acc = k,              x = phi + c k
//   kStageMid:   acc = acc + 2 k,      x = phi + c k
//   kStageLast:  x = phi + (acc + k) c (c = step / 6)
// x is updated in place: the stencil of entry j reads the original x[j-2], x[j-1]
// from a two-entry window and x[j+1], x[j+2] before they are overwritten.
template <int E, int P, int MODE>
__device__ inline void rk_stage(double (&xr)[E], double (&xi)[E], const double (&pr)[E], const double (&pi)[E],
                                double (&ar)[E], double (&ai)[E], double Ed, double c, const double *lds,
                                const Halo &h, int lane, int ll)
{
    // An offset the compiler cannot see through: the band reads stay inside the
    // stage instead of being hoisted out of the step loop into registers.
    int lo = ll;
    asm volatile("" : "+v"(lo));
    // entries lE-2, lE-1 (left) and lE+E, lE+E+1 (right); with one entry per lane
    // the second neighbours are two lanes away
    const double l1r = shfl_up1(xr[E - 1], lane), l1i = shfl_up1(xi[E - 1], lane);
    const double r1r = shfl_down1(xr[0], lane), r1i = shfl_down1(xi[0], lane);
    double l2r, l2i, r2r, r2i;
    if constexpr (E == 1) {
        l2r = shfl_up1(l1r, lane);
        l2i = shfl_up1(l1i, lane);
        r2r = shfl_down1(r1r, lane);
        r2i = shfl_down1(r1i, lane);
    } else {
        l2r = shfl_up1(xr[E - 2], lane);
        l2i = shfl_up1(xi[E - 2], lane);
        r2r = shfl_down1(xr[1], lane);
        r2i = shfl_down1(xi[1], lane);
    }
    double wr2 = l2r, wi2 = l2i, wr1 = l1r, wi1 = l1i;
    double c1re = h.o1re, c1im = h.o1im, c2a = h.o2a, c2b = h.o2b;
#pragma unroll
    for (int j = 0; j < E; j++) {
        const double cr = xr[j], ci = xi[j];
        const double p1r = j + 1 < E ? xr[j + 1] : r1r;
        const double p1i = j + 1 < E ? xi[j + 1] : r1i;
        const double p2r = j + 2 < E ? xr[j + 2] : (j + 2 == E ? r1r : r2r);
        const double p2i = j + 2 < E ? xi[j + 2] : (j + 2 == E ? r1i : r2i);
        const double dre = lds[(kRkDre * E + j) * P + lo];
        const double dim = lds[(kRkDim * E + j) * P + lo];
        const double o1re = lds[(kRkO1re * E + j) * P + lo];
        const double o1im = lds[(kRkO1im * E + j) * P + lo];
        const double o2 = lds[(kRkO2im * E + j) * P + lo];
        const double drv = lds[(kRkDrv * E + j) * P + lo];
        double re = 0.0, im = 0.0;
        re = __builtin_fma(-c2a, wi2, re);          // column i-2 (purely imaginary entry)
        im = __builtin_fma(c2a, wr2, im);
        re = __builtin_fma(c1re, wr1, re);          // column i-1
        re = __builtin_fma(-c1im, wi1, re);
        im = __builtin_fma(c1re, wi1, im);
        im = __builtin_fma(c1im, wr1, im);
        re = __builtin_fma(dre, cr, re);            // column i
        re = __builtin_fma(-dim, ci, re);
        im = __builtin_fma(dre, ci, im);
        im = __builtin_fma(dim, cr, im);
        re = __builtin_fma(o1re, p1r, re);          // column i+1
        re = __builtin_fma(-o1im, p1i, re);
        im = __builtin_fma(o1re, p1i, im);
        im = __builtin_fma(o1im, p1r, im);
        re = __builtin_fma(-o2, p2i, re);           // column i+2
        im = __builtin_fma(o2, p2r, im);
        const double dr = drv * cr, di = drv * ci;
        const double kr = __builtin_fma(Ed, di, im);
        const double ki = -__builtin_fma(Ed, dr, re);
        if (MODE == kStageFirst) {
            ar[j] = kr;
            ai[j] = ki;
            xr[j] = __builtin_fma(c, kr, pr[j]);
            xi[j] = __builtin_fma(c, ki, pi[j]);
        } else if (MODE == kStageMid) {
            ar[j] = __builtin_fma(2.0, kr, ar[j]);
            ai[j] = __builtin_fma(2.0, ki, ai[j]);
            xr[j] = __builtin_fma(c, kr, pr[j]);
            xi[j] = __builtin_fma(c, ki, pi[j]);
        } else {
            xr[j] = __builtin_fma(ar[j] + kr, c, pr[j]);
            xi[j] = __builtin_fma(ai[j] + ki, c, pi[j]);
        }
        wr2 = wr1;
        wi2 = wi1;
        wr1 = cr;
        wi1 = ci;
        c2a = c2b;
        c2b = o2;
        c1re = o1re;
        c1im = o1im;
    }
}

// E(t) at the three stage times of an integration from t: t, t + step/2 (stages 2
// and 3) and t + step, with t advanced by step * 0.5 twice as rk_int_dimer does.
__device__ inline void rk_drives(const RkArgs &a, double &t, double step, double &e0, double &e1, double &e2)
{
    e0 = e1 = e2 = 0.0;
#pragma unroll 1
    for (int q = 0; q < 3; q++) {      // one drive evaluation site
        const double v = rk_drive(a, t);
        e0 = q == 0 ? v : e0;
        e1 = q == 1 ? v : e1;
        e2 = q == 2 ? v : e2;
        if (q < 2) t = t + step * 0.5;
    }
}

// The drives of the first (full rk_step) integration of each step of a launch:
// they depend on the step index only, so one thread per step writes them and the
// trajectories read them (crossing sub-steps evaluate their own).
__global__ void rk_drive_table_kernel(RkArgs a)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.nsteps) return;
    const long long gs = a.pos0 + k;
    const long long period_id = gs / a.rk_ns, step_id = gs % a.rk_ns;
    double t = double(period_id) * a.T + double(step_id) * a.rk_step;
    double e0, e1, e2;
    rk_drives(a, t, a.rk_step, e0, e1, e2);
    a.tab[3 * k] = e0;
    a.tab[3 * k + 1] = e1;
    a.tab[3 * k + 2] = e2;
}

// rk_int_dimer: x = RK4(phi, step) with stage drives e0, e1, e2.  The stage
// arguments use the global rk_step whatever step is (arg_upd reads ed->rk_step,
// rk_proc.cpp:150-173).
template <int E, int P>
__device__ inline void rk_integrate(const RkArgs &a, double step, double e0, double e1, double e2,
                                    const double (&pr)[E], const double (&pi)[E], double (&xr)[E],
                                    double (&xi)[E], double (&ar)[E], double (&ai)[E], const double *lds,
                                    const Halo &h, int lane, int ll)
{
    const double hh = 0.5 * a.rk_step, H = 1.0 * a.rk_step;
#pragma unroll
    for (int j = 0; j < E; j++) {
        xr[j] = pr[j];
        xi[j] = pi[j];
    }
    rk_stage<E, P, kStageFirst>(xr, xi, pr, pi, ar, ai, e0, hh, lds, h, lane, ll);
    rk_stage<E, P, kStageMid>(xr, xi, pr, pi, ar, ai, e1, hh, lds, h, lane, ll);
    rk_stage<E, P, kStageMid>(xr, xi, pr, pi, ar, ai, e1, H, lds, h, lane, ll);
    rk_stage<E, P, kStageLast>(xr, xi, pr, pi, ar, ai, e2, step / 6.0, lds, h, lane, ll);
}

// Waves per SIMD the register allocation aims for at each lane width.
template <int E>
struct RkOcc {
    static constexpr int waves = E <= 2 ? 4 : E == 4 ? 3 : E == 8 ? 2 : 1;
};

template <int E, int P>
__global__ __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(RkOcc<E>::waves)))
void rk_kernel(RkArgs a)
{
    extern __shared__ double rk_lds[];
    constexpr int NB = kRkStageBands * E * P;
    for (int i = threadIdx.x; i < NB; i += 64 * kWaves) rk_lds[i] = a.bands[i];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int ll = P == 64 ? lane : lane & (P - 1);      // lane within the trajectory's group
    const int tr = P == 64 ? blockIdx.x * kWaves + (int)(threadIdx.x >> 6)
                           : (blockIdx.x * kWaves + (int)(threadIdx.x >> 6)) * (64 / P) + lane / P;
    if (tr >= a.M) return;

    TrajState s = a.st[tr];
    double *col = (double *)(s.cur ? a.slot1 : a.slot0) + (size_t)tr * a.L;
    double pr[E], pi[E], xr[E], xi[E], ar[E], ai[E];
    // dp = 64 E for an RK context: every entry is inside the column (the padding
    // beyond d is zero and stays zero), so the loads and stores need no guard
#pragma unroll
    for (int j = 0; j < E; j++) {
        pr[j] = col[ll * E + j];
        pi[j] = col[a.dp + ll * E + j];
    }
    Halo h;
    // zero at the group's first lane: those columns lie before entry 0
    h.o1re = ll ? rk_lds[(kRkO1re * E + E - 1) * P + ll - 1] : 0.0;
    h.o1im = ll ? rk_lds[(kRkO1im * E + E - 1) * P + ll - 1] : 0.0;
    // H_part(lE-2, lE): entry E-2 of lane l-1, or with one entry per lane lane l-2's
    if constexpr (E == 1)
        h.o2a = ll >= 2 ? rk_lds[kRkO2im * P + ll - 2] : 0.0;
    else
        h.o2a = ll ? rk_lds[(kRkO2im * E + E - 2) * P + ll - 1] : 0.0;
    h.o2b = ll ? rk_lds[(kRkO2im * E + E - 1) * P + ll - 1] : 0.0;
    const double *Ld = a.bands + (size_t)kRkLd * E * P;
    const double *Lm = a.bands + (size_t)kRkLm * E * P;
    const double *Lp = a.bands + (size_t)kRkLp * E * P;

    unsigned long long integrations = 0, jumps = 0;
    bool failed = false;
    double pn = group_norm<E, P>(pr, pi);       // ||phi||^2 of the current state
    double t = s.t;
    for (int k = 0; k < a.nsteps && !failed; k++) {
        const long long gs = a.pos0 + k;
        const long long period_id = gs / a.rk_ns, step_id = gs % a.rk_ns;
        t = double(period_id) * a.T + double(step_id) * a.rk_step;   // rk_period
        double step = a.rk_step, estep = 0.0;
        bool partial = false;   // integrating the part of a crossed step before its jump
        int crossings = 0;
        for (;;) {   // rk_step_dimer, its tail recursion as a loop (one integration site)
            double e0, e1, e2;
            if (!partial && crossings == 0) {   // the step's full integration: tabulated drives
                const int kk = __builtin_amdgcn_readfirstlane(k);
                e0 = a.tab[3 * kk];
                e1 = a.tab[3 * kk + 1];
                e2 = a.tab[3 * kk + 2];
                t = t + step * 0.5;
                t = t + step * 0.5;
            } else {
                rk_drives(a, t, step, e0, e1, e2);
            }
            rk_integrate<E, P>(a, step, e0, e1, e2, pr, pi, xr, xi, ar, ai, rk_lds, h, lane, ll);
            integrations++;
            if (!partial) {
                const double n = group_norm<E, P>(xr, xi);
                // is_norm_crossed throws for ||phi||^2 > 1 (experiment.cpp:1247-1251); a
                // NaN norm or an endless run of crossings (the reference would recurse
                // without end) stops the trajectory with error code 2
                const int code = n > 1.0 ? 1 : (!(n == n) || crossings >= kRkMaxCrossings) ? 2 : 0;
                if (code) {
                    if (ll == 0 && atomicCAS(&a.err[0], 0, code) == 0) {
                        const unsigned long long bits = __double_as_longlong(n);
                        a.err[1] = tr;
                        a.err[2] = (int)(bits & 0xffffffffULL);
                        a.err[3] = (int)(bits >> 32);
                    }
                    failed = true;
                    break;
                }
                if (n > s.eta) {            // no crossing: accept
#pragma unroll
                    for (int j = 0; j < E; j++) {
                        pr[j] = xr[j];
                        pi[j] = xi[j];
                    }
                    pn = n;
                    break;
                }
                // crossing: linear interpolation of ||phi||^2, restore the step-start
                // state and time, integrate the first part (next pass), jump, go on
                const double diff = pn - n;
                const double bp = (pn - s.eta) / diff;
                const double ep = 1.0 - bp;
                const double bstep = step * bp;
                estep = step * ep;
                t = t - step;
                step = bstep;
                partial = true;
                crossings++;
                continue;
            }
#pragma unroll
            for (int j = 0; j < E; j++) {
                pr[j] = xr[j];
                pi[j] = xi[j];
            }
            const double jn = group_norm<E, P>(pr, pi);
            if (a.jump > 0 && a.is_obs == 1) {
                if (ll == 0 && s.logn < a.log.cap) {
                    const size_t o = (size_t)tr * a.log.cap + s.logn;
                    a.log.time[o] = t;
                    a.log.norm[o] = jn;
                    a.log.eta[o] = s.eta;
                    a.log.type[o] = 0;
                }
                s.logn++;
                s.njumps++;
            }
            // rk_recovery (one dissipator): normalise, draw ran, phi = L phi / ||L phi||
            const double nrm = sqrt(jn);
#pragma unroll
            for (int j = 0; j < E; j++) {
                pr[j] = pr[j] / nrm;
                pi[j] = pi[j] / nrm;
            }
            (void)mcg_next(s.rng, a.stride);
            int lo = ll;
            asm volatile("" : "+v"(lo));
            const double lr = shfl_up1(pr[E - 1], lane), li = shfl_up1(pi[E - 1], lane);
            const double rr = shfl_down1(pr[0], lane), ri = shfl_down1(pi[0], lane);
#pragma unroll
            for (int j = 0; j < E; j++) {
                const double mr = j ? pr[j - 1] : lr, mi = j ? pi[j - 1] : li;
                const double qr = j + 1 < E ? pr[j + 1] : rr, qi = j + 1 < E ? pi[j + 1] : ri;
                const double cm = Lm[j * P + lo], cd = Ld[j * P + lo], cp = Lp[j * P + lo];
                double re = 0.0, im = 0.0;
                re = __builtin_fma(cm, mr, re);
                im = __builtin_fma(cm, mi, im);
                re = __builtin_fma(cd, pr[j], re);
                im = __builtin_fma(cd, pi[j], im);
                re = __builtin_fma(cp, qr, re);
                im = __builtin_fma(cp, qi, im);
                ar[j] = re;
                ai[j] = im;
            }
            const double g = group_norm<E, P>(ar, ai) * 1.0;
            const double sg = sqrt(g / 1.0);
#pragma unroll
            for (int j = 0; j < E; j++) {
                pr[j] = ar[j] / sg;
                pi[j] = ai[j] / sg;
            }
            jumps++;
            double ev = mcg_next(s.rng, a.stride);
            while (ev == 0.0) ev = mcg_next(s.rng, a.stride);
            s.eta = ev;
            pn = group_norm<E, P>(pr, pi);
            step = estep;
            partial = false;
        }
    }
    s.t = t;
#pragma unroll
    for (int j = 0; j < E; j++) {
        col[ll * E + j] = pr[j];
        col[a.dp + ll * E + j] = pi[j];
    }
    if (ll == 0) {
        a.st[tr] = s;
        atomicAdd(&a.counters[0], 4ULL * integrations);
        atomicAdd(&a.counters[1], jumps);
    }
}

template <int E, int P>
hipError_t launch_rk_e(const RkArgs &a, hipStream_t s)
{
    const size_t lds = (size_t)kRkStageBands * E * P * sizeof(double);
    if (lds > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&rk_kernel<E, P>),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    rk_drive_table_kernel<<<(a.nsteps + 255) / 256, 256, 0, s>>>(a);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    constexpr int per_block = kWaves * (64 / P);   // trajectories per workgroup
    const int blocks = (a.M + per_block - 1) / per_block;
    rk_kernel<E, P><<<blocks, 64 * kWaves, lds, s>>>(a);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_rk(const RkArgs &a, hipStream_t s)
{
    if (a.M <= 0 || a.nsteps <= 0) return hipSuccess;
    if (a.dp != rk_group(a.d) * rk_lanes(a.d) || a.L != 2 * a.dp) return hipErrorInvalidValue;
    if (rk_group(a.d) == 32) return rk_lanes(a.d) == 4 ? launch_rk_e<4, 32>(a, s) : hipErrorInvalidValue;
    switch (rk_lanes(a.d)) {
    case 1: return launch_rk_e<1, 64>(a, s);
    case 2: return launch_rk_e<2, 64>(a, s);
    case 4: return launch_rk_e<4, 64>(a, s);
    case 8: return launch_rk_e<8, 64>(a, s);
    case 16: return launch_rk_e<16, 64>(a, s);
    case 32: return launch_rk_e<32, 64>(a, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace qjx
