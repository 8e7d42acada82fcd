// qjx_lindblad.hip -- the dimer's Lindblad equation integrated in rho space on the GPU
// (FP64), the ensemble-level counterpart of the trajectory engine.
//
//   d rho / dt = -i [H(t), rho] + g (L rho L^dag - 1/2 {L^dag L, rho}),  g = gamma / N,
//   H(t) = H + E(t) H_drv,  E(t) = E + A on the first half period, E - A on the second.
//
// The reference integrates this equation with RK4 in the SU(N) f-basis
// (CQdiss_os_dimer/CQdiss/calcODE.cpp:310-389 calcODE_std; rho recovered by
// CalcRho.cpp:4-67).  A fixed linear change of variables commutes with RK4, so RK4
// on rho itself is the same scheme.  H and L are tridiagonal and L^dag L
// pentadiagonal, so a right-hand side is a stencil of radius 2 in the indices of
// rho.  One kernel performs a whole RK4 step per tile of rho (its four stages in LDS
// on overlapping windows), so a step reads rho once and writes it once, in one
// launch.
//
// The step kernel evaluates the right-hand side with exactly the operations, in
// exactly the order, of the CPU restatement used as the oracle, with floating-point
// contraction off, so both produce the same bits.
//
// The drive E(t) follows CQdiss's calcVectValue_real (calcODE.cpp:100-116): at every
// stage time t (time, time + h/2, time + h/2, time + h with time = p T + s h) it is
// E + A sin(omega t + phase) (dimer_drv_type 1) or E + A sign(sin(omega t + phase))
// (type 0, a zero sine counting as positive), so at each switch of the square drive
// one stage takes the neighbouring half period's value, as there.  The sine is the
// Cody-Waite + Taylor code the CPU restatement uses.  The other rule holds the
// square drive constant over each step, steps aligned with the half periods.
// The step kernel reads the step counter from device memory, so one captured graph
// holds many consecutive steps.
#include <hip/hip_runtime.h>

#include <cmath>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "../../include/qjx.h"
#include "qjx_internal.h"
#include "qjx_kernels.h"
#include "qjx_runner.h"

#pragma clang fp contract(off)

namespace qjx {

// band matrices of the model on the device: H diagonal and off-diagonal, H_drv
// diagonal, L (L_kk = ld, L_{k+1,k} = -lo, L_{k,k+1} = lo), L^dag L bands q0..q2
struct LBands {
    const double *hd, *hdrv, *ho, *ld, *lo, *q0, *q1, *q2;
    double g;
    int d;
};

// E(t): the drive at the stage times of step g = *step + j (qjx_lindblad_drive)
struct LDrive {
    const long long *step;    // steps done since the integration's t = 0
    double T, h, E0, A, freq, phase, Eplus, Eminus;
    int S, type, per_stage;
};

namespace {

typedef double2 c2;

__device__ inline c2 cadd(c2 a, c2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline c2 csub(c2 a, c2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline c2 rmul(double s, c2 a) { return make_double2(s * a.x, s * a.y); }
__device__ inline c2 mulr(c2 a, double s) { return make_double2(a.x * s, a.y * s); }

// The band coefficients of a window of w row or column indices, held in LDS: band
// b's entry k at s[b * w + (k - k0)] (b: hd, hdrv, ho, ld, lo, q0, q1, q2).
constexpr int kBands = 8;
struct BandView {
    const double *s;
    int k0, w;
    __device__ double at(int b, int k) const { return s[b * w + (k - k0)]; }
    __device__ double hd(int k) const { return at(0, k); }
    __device__ double hdrv(int k) const { return at(1, k); }
    __device__ double ho(int k) const { return at(2, k); }
    __device__ double ld(int k) const { return at(3, k); }
    __device__ double lo(int k) const { return at(4, k); }
    __device__ double q0(int k) const { return at(5, k); }
    __device__ double q1(int k) const { return at(6, k); }
    __device__ double q2(int k) const { return at(7, k); }
};

// (L x)_i over the rows of column j (L_ii = ld, L_{i,i-1} = -lo[i-1], L_{i,i+1} = lo[i]);
// x(r, c) returns the argument's entry (r, c), R the bands at row indices
template <class X, class BR>
__device__ inline c2 lrow(const BR &R, int d, const X &x, int i, int j)
{
    c2 s = rmul(R.ld(i), x(i, j));
    if (i > 0) s = cadd(s, rmul(-R.lo(i - 1), x(i - 1, j)));
    if (i + 1 < d) s = cadd(s, rmul(R.lo(i), x(i + 1, j)));
    return s;
}

// right-hand side at (i, j); terms in the order of the CPU restatement.  Reads the
// argument at (i +- 2, j), (i, j +- 2) and (i +- 1, j +- 1), the bands at rows
// i - 2 .. i and columns j - 2 .. j, inside 0..d-1 only.  R: the bands at row
// indices, C: at column indices (the same values), g the dissipation rate.
template <class X, class BR, class BC>
__device__ inline c2 rhs_at(const BR &R, const BC &C, int d, double g, double E, const X &x, int i, int j)
{
    // (L x L^T)_ij from T = L x: T_ij, T_i,j-1, T_i,j+1
    c2 lrl = rmul(C.ld(j), lrow(R, d, x, i, j));
    if (j > 0) lrl = cadd(lrl, rmul(-C.lo(j - 1), lrow(R, d, x, i, j - 1)));
    if (j + 1 < d) lrl = cadd(lrl, rmul(C.lo(j), lrow(R, d, x, i, j + 1)));
    const double hdi = R.hd(i) + E * R.hdrv(i), hdj = C.hd(j) + E * C.hdrv(j);
    const c2 xij = x(i, j);
    c2 hr = rmul(hdi, xij);
    if (i > 0) hr = cadd(hr, rmul(R.ho(i - 1), x(i - 1, j)));
    if (i + 1 < d) hr = cadd(hr, rmul(R.ho(i), x(i + 1, j)));
    c2 rh = mulr(xij, hdj);
    if (j > 0) rh = cadd(rh, mulr(x(i, j - 1), C.ho(j - 1)));
    if (j + 1 < d) rh = cadd(rh, mulr(x(i, j + 1), C.ho(j)));
    c2 qr = rmul(R.q0(i), xij);
    if (i + 1 < d) qr = cadd(qr, rmul(R.q1(i), x(i + 1, j)));
    if (i + 2 < d) qr = cadd(qr, rmul(R.q2(i), x(i + 2, j)));
    if (i >= 1) qr = cadd(qr, rmul(R.q1(i - 1), x(i - 1, j)));
    if (i >= 2) qr = cadd(qr, rmul(R.q2(i - 2), x(i - 2, j)));
    c2 rq = mulr(xij, C.q0(j));
    if (j + 1 < d) rq = cadd(rq, mulr(x(i, j + 1), C.q1(j)));
    if (j + 2 < d) rq = cadd(rq, mulr(x(i, j + 2), C.q2(j)));
    if (j >= 1) rq = cadd(rq, mulr(x(i, j - 1), C.q1(j - 1)));
    if (j >= 2) rq = cadd(rq, mulr(x(i, j - 2), C.q2(j - 2)));
    // (0 - i) * (hr - rh): the complex product with exact zeros, as std::complex
    const c2 c = csub(hr, rh);
    const c2 mi = make_double2(0.0 * c.x - (-1.0) * c.y, 0.0 * c.y + (-1.0) * c.x);
    const c2 diss = rmul(g, csub(lrl, rmul(0.5, cadd(qr, rq))));
    return cadd(mi, diss);
}

// sin x: quadrant k = nearest integer to x 2/pi, r = x - k pi/2 with pi/2 split into
// 33 + 33 + 53 significant bits (Cody & Waite 1980), then the Taylor series of sin r
// or cos r on |r| <= pi/4 by Horner: the operations of the CPU restatement.
__device__ inline double lb_sin(double x)
{
    const double k = rint(x * 0x1.45f306dc9c883p-1);
    double r = x - k * 0x1.921fb54400000p+0;
    r = r - k * 0x1.0b4611a600000p-34;
    r = r - k * 0x1.3198a2e037073p-69;
    const double z = r * r;
    double ps = -0x1.761b41316381ap-75;
    ps = fma(ps, z, 0x1.71b8ef6dcf572p-66);
    ps = fma(ps, z, -0x1.2f49b46814157p-57);
    ps = fma(ps, z, 0x1.952c77030ad4ap-49);
    ps = fma(ps, z, -0x1.ae7f3e733b81fp-41);
    ps = fma(ps, z, 0x1.6124613a86d09p-33);
    ps = fma(ps, z, -0x1.ae64567f544e4p-26);
    ps = fma(ps, z, 0x1.71de3a556c734p-19);
    ps = fma(ps, z, -0x1.a01a01a01a01ap-13);
    ps = fma(ps, z, 0x1.1111111111111p-7);
    ps = fma(ps, z, -0x1.5555555555555p-3);
    double pc = 0x1.f2cf01972f578p-80;
    pc = fma(pc, z, -0x1.0ce396db7f853p-70);
    pc = fma(pc, z, 0x1.e542ba4020225p-62);
    pc = fma(pc, z, -0x1.6827863b97d97p-53);
    pc = fma(pc, z, 0x1.ae7f3e733b81fp-45);
    pc = fma(pc, z, -0x1.93974a8c07c9dp-37);
    pc = fma(pc, z, 0x1.1eed8eff8d898p-29);
    pc = fma(pc, z, -0x1.27e4fb7789f5cp-22);
    pc = fma(pc, z, 0x1.a01a01a01a01ap-16);
    pc = fma(pc, z, -0x1.6c16c16c16c17p-10);
    pc = fma(pc, z, 0x1.5555555555555p-5);
    pc = fma(pc, z, -0x1.0000000000000p-1);
    const double s = fma(r * z, ps, r);
    const double c = fma(z, pc, 1.0);
    const int q = (int)((long long)k & 3);
    return q == 0 ? s : q == 1 ? c : q == 2 ? -s : -c;
}

// E at stage `stage` (0..3) of step *D.step + j
__device__ inline double stage_drive(const LDrive &D, int j, int stage)
{
    const long long g = *D.step + j;
    const long long p = g / D.S, s = g % D.S;
    if (!D.per_stage) return s < D.S / 2 ? D.Eplus : D.Eminus;
    const double time = double(p) * D.T + double(s) * D.h;
    const double t = stage == 0 ? time : stage == 3 ? time + D.h : time + D.h / 2.0;
    const double sn = lb_sin(D.freq * t + D.phase);
    const double f = D.type == 1 ? sn : (sn < 0.0 ? -1.0 : 1.0);
    return D.E0 + D.A * f;
}

// K RK4 steps (K = 1 or 2), all their stages in one launch: workgroup (ti, tj) owns
// the T x T tile of rho at (i0, j0) = (T ti, T tj).  A stage's right-hand side
// reads its argument up to two entries away, so a step shrinks the window it can
// evaluate by 8 on every side: the launch reads rho on the tile widened by 8K, and
// step s (0..K-1) evaluates its stages on the window widened by 8(K - s) - 2,
// - 4, - 6 and its new rho on the window widened by 8(K - 1 - s), over the old rho
// in LDS (each entry's update reads the old rho at that entry only).  Everything
// between the read of rho and the write of the new tile stays in LDS, the band
// coefficients of the window's rows and columns included; the neighbouring tiles'
// values are recomputed, not exchanged (2x the stencil work at T = 16 and K = 1,
// 3.4x at T = 8 and K = 1, which gives small d four times the workgroups).  Per
// entry and step:
//   k1 = f(rho), v1 = rho + (h/2) k1;  k2 = f(v1), v2 = rho + (h/2) k2;
//   k3 = f(v2), v3 = rho + h k3;       k4 = f(v3);
//   rho' = rho + (h/6) (((k1 + 2 k2) + 2 k3) + k4)
// the operations of the CPU restatement in its order.  rho is read from `in` and
// the tile written to `out` (other workgroups still read `in`'s halo).  NT threads.
// a stage argument held in LDS over a square window: entry (r, c) at
// buf[(r - r0) * w + (c - c0)]
struct LdsView {
    c2 *buf;
    int r0, c0, w;
    __device__ c2 &operator()(int r, int c) const { return buf[(r - r0) * w + (c - c0)]; }
};

template <int T, int NT, int K>
__global__ __launch_bounds__(NT) void rk4_step_kernel(LBands B, LDrive D, int j, const c2 *in, c2 *out, double h)
{
    constexpr int H = 8 * K;                 // halo of the launch
    constexpr int W0 = T + 2 * H;            // rho and the bands
    constexpr int WA = W0 - 4;               // v1 (and v3) of the first step
    constexpr int WB = W0 - 8;               // v2 of the first step
    constexpr int WF = W0 - 16;              // the first step's new rho and combination
    static_assert(2 * kBands * W0 <= NT, "band load");
    __shared__ c2 s_rho[W0 * W0];
    __shared__ c2 s_a[WA * WA];
    __shared__ c2 s_b[WB * WB];
    __shared__ c2 s_acc[WF * WF];
    __shared__ double s_brow[kBands * W0], s_bcol[kBands * W0];
    const int d = B.d, t = threadIdx.x;
    const int i0 = blockIdx.y * T, j0 = blockIdx.x * T;
    const double c2h = 0.5 * h, h6 = h / 6.0;
    // rho on the widened tile (entries outside 0..d-1 are never read)
    for (int q = t; q < W0 * W0; q += NT) {
        const int r = i0 - H + q / W0, c = j0 - H + q % W0;
        s_rho[q] = (r >= 0 && r < d && c >= 0 && c < d) ? in[(size_t)r * d + c] : make_double2(0.0, 0.0);
    }
    // the bands at the window's row and column indices
    if (t < 2 * kBands * W0) {
        const int col = t >= kBands * W0, u = t % (kBands * W0);
        const int b = u / W0, k = (col ? j0 : i0) - H + u % W0;
        const double *src[kBands] = {B.hd, B.hdrv, B.ho, B.ld, B.lo, B.q0, B.q1, B.q2};
        (col ? s_bcol : s_brow)[u] = (k >= 0 && k < d) ? src[b][k] : 0.0;
    }
    const BandView R{s_brow, i0 - H, W0}, C{s_bcol, j0 - H, W0};
    const LdsView rho{s_rho, i0 - H, j0 - H, W0};
    double E[K][4];
#pragma unroll
    for (int s = 0; s < K; s++)
#pragma unroll
        for (int st = 0; st < 4; st++) E[s][st] = stage_drive(D, j + s, st);
    __syncthreads();
#pragma unroll
    for (int s = 0; s < K; s++) {
        const int g = 8 * (K - s);           // this step's input window: the tile widened by g
        const int wf = T + 2 * (g - 8);      // its output window
        const LdsView acc{s_acc, i0 - (g - 8), j0 - (g - 8), wf};
        // stages 0..2: v = rho + c k over the tile widened by w, and the running
        // combination on the output window
        auto stage = [&](const LdsView &x, double Es, double c, const LdsView &v, int w, int st) {
            const int n = T + 2 * w;
            for (int q = t; q < n * n; q += NT) {
                const int r = i0 - w + q / n, cc = j0 - w + q % n;
                if (r < 0 || r >= d || cc < 0 || cc >= d) continue;
                const c2 k = rhs_at(R, C, d, B.g, Es, x, r, cc);
                v(r, cc) = cadd(rho(r, cc), rmul(c, k));
                const int ai = r - (i0 - (g - 8)), aj = cc - (j0 - (g - 8));
                if (ai >= 0 && ai < wf && aj >= 0 && aj < wf) {
                    c2 &a = acc(r, cc);
                    a = st == 0 ? k : cadd(a, rmul(2.0, k));
                }
            }
        };
        const LdsView v1{s_a, i0 - (g - 2), j0 - (g - 2), T + 2 * (g - 2)};
        const LdsView v2{s_b, i0 - (g - 4), j0 - (g - 4), T + 2 * (g - 4)};
        const LdsView v3{s_a, i0 - (g - 6), j0 - (g - 6), T + 2 * (g - 6)};
        stage(rho, E[s][0], c2h, v1, g - 2, 0);
        __syncthreads();
        stage(v1, E[s][1], c2h, v2, g - 4, 1);
        __syncthreads();
        stage(v2, E[s][2], h, v3, g - 6, 2);     // v3 over v1's buffer (v1 is no longer read)
        __syncthreads();
        const int w = g - 8;
        for (int q = t; q < wf * wf; q += NT) {
            const int r = i0 - w + q / wf, cc = j0 - w + q % wf;
            if (r < 0 || r >= d || cc < 0 || cc >= d) continue;
            const c2 k = rhs_at(R, C, d, B.g, E[s][3], v3, r, cc);
            const c2 sum = cadd(acc(r, cc), k);
            const c2 nr = cadd(rho(r, cc), rmul(h6, sum));
            if (s == K - 1) out[(size_t)r * d + cc] = nr;
            else rho(r, cc) = nr;
        }
        if (s < K - 1) __syncthreads();
    }
}

// the step counter advances by n (one thread)
__global__ void lb_advance_kernel(long long *step, int n) { *step += n; }

// <k>, <(k - <k>)^2>, <H + E0 H_drv>, adr_k = rho_kk (fixed-order sums, one workgroup)
__global__ __launch_bounds__(256) void lb_obs_kernel(LBands B, double E0, const c2 *rho, double *out, double *adr)
{
    __shared__ double s[3][256];
    const int d = B.d, t = threadIdx.x;
    double mean = 0.0, en = 0.0;
    for (int k = t; k < d; k += 256) {
        const double p = rho[(size_t)k * d + k].x;
        if (adr) adr[k] = p;
        mean += (double)k * p;
        en += (B.hd[k] + E0 * B.hdrv[k]) * p;
        if (k + 1 < d) en += B.ho[k] * (rho[(size_t)k * d + k + 1].x + rho[(size_t)(k + 1) * d + k].x);
    }
    s[0][t] = mean;
    s[1][t] = en;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (t < off) {
            s[0][t] += s[0][t + off];
            s[1][t] += s[1][t + off];
        }
        __syncthreads();
    }
    const double m = s[0][0];
    __syncthreads();
    double m2 = 0.0;
    for (int k = t; k < d; k += 256) {
        const double p = rho[(size_t)k * d + k].x;
        m2 += ((double)k - m) * ((double)k - m) * p;
    }
    s[2][t] = m2;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (t < off) s[2][t] += s[2][t + off];
        __syncthreads();
    }
    if (t == 0) {
        out[0] = s[0][0];
        out[1] = s[2][0];
        out[2] = s[1][0];
    }
}

}  // namespace
}  // namespace qjx

using qjx::set_last_error;

struct qjx_lindblad {
    int d = 0, steps = 0;
    double T = 0.0, h = 0.0, Estat = 0.0;
    hipStream_t stream = nullptr;
    double *bands = nullptr;          // hd, hdrv, ho, ld, lo, q0, q1, q2 (d each)
    double2 *buf = nullptr;           // rho and the other buffer of a step (d x d each)
    double *obs = nullptr;            // 3 observables + purity + adr[d]
    long long *d_step = nullptr;      // RK4 steps since t = 0 (device)
    qjx::LBands B{};
    qjx::LDrive D{};
    bool graphs = true;
    int tile = 16;                    // step kernel tile edge: 8 or 16
    int fuse = 1;                     // RK4 steps per launch with 8 x 8 tiles: 1 or 2
    hipGraphExec_t exec = nullptr;    // kChunk consecutive RK4 steps and the counter advance
    hipGraph_t graph = nullptr;
};

namespace {
constexpr int kChunk = 48;            // RK4 steps per graph launch
static_assert(kChunk % 4 == 0, "a chunk has an even number of launches (one or two steps each)");
}

namespace {

int lb_fail(int code, const std::string &msg)
{
    set_last_error(msg);
    return code;
}

#define LB_TRY(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return lb_fail(e_ == hipErrorOutOfMemory ? QJX_ERR_NOMEM : QJX_ERR_HIP,           \
                           std::string(#expr) + ": " + hipGetErrorString(e_));                \
    } while (0)

size_t nsq(const qjx_lindblad *l) { return (size_t)l->d * l->d; }
double2 *rho_of(qjx_lindblad *l) { return l->buf; }

// the band matrices of H, H_drv, L and L^dag L (reference newdel.cpp:450-498,
// 1454-1498), summed as the CPU restatement sums them
void build_bands(const qjx_model_desc *m, int d, std::vector<double> &h)
{
    h.assign((size_t)8 * d, 0.0);
    double *hd = h.data(), *hdrv = hd + d, *ho = hd + 2 * d, *ld = hd + 3 * d, *lo = hd + 4 * d;
    double *q0 = hd + 5 * d, *q1 = hd + 6 * d, *q2 = hd + 7 * d;
    const double U = m->prm_U / double(d - 1);
    for (int k = 0; k < d; k++) {
        hd[k] = 2.0 * U * double(k * (k - 1) + (d - (k + 1)) * (d - (k + 1) - 1));
        hdrv[k] = double((d - (k + 1)) - k);
        ld[k] = double(k - (d - (k + 1)));
    }
    for (int k = 0; k + 1 < d; k++) {
        ho[k] = -m->prm_J * std::sqrt(double((d - (k + 1)) * (k + 1)));
        lo[k] = std::sqrt(double((d - (k + 1)) * (k + 1)));
    }
    auto Lel = [&](int i, int j) -> double {
        if (i == j) return ld[i];
        if (i == j + 1) return -lo[j];
        if (j == i + 1) return lo[i];
        return 0.0;
    };
    for (int i = 0; i < d; i++)
        for (int off = 0; off <= 2 && i + off < d; off++) {
            const int j = i + off;
            double s = 0.0;
            for (int k = 0; k < d; k++) s += Lel(k, i) * Lel(k, j);
            (off == 0 ? q0 : off == 1 ? q1 : q2)[i] = s;
        }
}

// n consecutive RK4 steps, then the counter advances by n.  The step kernel reads
// the drive's step from the counter.  Each launch writes rho to the other buffer;
// after an odd number of launches rho is copied back, so it always ends in its own
// buffer (a captured chunk has an even number of launches and no copy).
int launch_steps(qjx_lindblad *l, int n, hipStream_t s)
{
    const size_t sq = nsq(l);
    double2 *rho = l->buf, *alt = rho + sq;
    const int T = l->tile;
    const int tiles = (l->d + T - 1) / T;
    const dim3 grid(tiles, tiles);
    int launches = 0;
    for (int j = 0; j < n;) {
        const double2 *in = (launches & 1) ? alt : rho;
        double2 *out = (launches & 1) ? rho : alt;
        const int k = (T == 8 && l->fuse == 2 && n - j >= 2) ? 2 : 1;
        if (k == 2) qjx::rk4_step_kernel<8, 1024, 2><<<grid, 1024, 0, s>>>(l->B, l->D, j, in, out, l->h);
        else if (T == 8) qjx::rk4_step_kernel<8, 512, 1><<<grid, 512, 0, s>>>(l->B, l->D, j, in, out, l->h);
        else qjx::rk4_step_kernel<16, 1024, 1><<<grid, 1024, 0, s>>>(l->B, l->D, j, in, out, l->h);
        j += k;
        launches++;
    }
    if (launches & 1) LB_TRY(hipMemcpyAsync(rho, alt, sq * sizeof(double2), hipMemcpyDeviceToDevice, s));
    qjx::lb_advance_kernel<<<1, 1, 0, s>>>(l->d_step, n);
    return QJX_OK;
}

}  // namespace

extern "C" void qjx_lindblad_destroy(qjx_lindblad *l)
{
    if (!l) return;
    if (l->stream) (void)hipStreamSynchronize(l->stream);
    if (l->exec) (void)hipGraphExecDestroy(l->exec);
    if (l->graph) (void)hipGraphDestroy(l->graph);
    if (l->d_step) (void)hipFree(l->d_step);
    if (l->bands) (void)hipFree(l->bands);
    if (l->buf) (void)hipFree(l->buf);
    if (l->obs) (void)hipFree(l->obs);
    if (l->stream) (void)hipStreamDestroy(l->stream);
    delete l;
}

extern "C" int qjx_lindblad_create(const qjx_model_desc *m, int steps_per_period, qjx_lindblad **out)
{
    const qjx_lindblad_drive aligned{0, 0.0, 0};
    return qjx_lindblad_create_drive(m, steps_per_period, &aligned, out);
}

extern "C" int qjx_lindblad_create_drive(const qjx_model_desc *m, int steps_per_period,
                                         const qjx_lindblad_drive *drive, qjx_lindblad **out)
{
    if (!m || !drive || !out) return lb_fail(QJX_ERR_ARG, "null argument");
    *out = nullptr;
    if (m->N < 1 || m->N + 1 > 4096) return lb_fail(QJX_ERR_ARG, "N must be in [1, 4095]");
    if (steps_per_period < 2 || (steps_per_period % 2)) return lb_fail(QJX_ERR_ARG, "steps_per_period must be even and >= 2");
    if (!(m->drv_freq > 0.0)) return lb_fail(QJX_ERR_ARG, "dimer_drv_freq must be > 0");
    if (!drive->per_stage && (drive->drv_type == 1 || drive->drv_phase != 0.0))
        return lb_fail(QJX_ERR_ARG, "a drive held constant per step is the square drive without phase");
    qjx_lindblad *l = new qjx_lindblad;
    l->d = m->N + 1;
    l->steps = steps_per_period;
    l->T = 2.0 * 3.1415926535897932384626433832795 / m->drv_freq;
    l->h = l->T / double(steps_per_period);
    l->Estat = m->prm_E;
    if (const char *env = std::getenv("QJX_GRAPH"))
        if (std::string(env) == "0") l->graphs = false;
    // 8 x 8 tiles for small d (four times the workgroups of 16 x 16, for 1.7x the
    // recomputed stencil work), 16 x 16 above; QJX_LB_TILE=8|16 forces one
    l->tile = l->d <= 160 ? 8 : 16;
    if (const char *env = std::getenv("QJX_LB_TILE")) {
        const std::string v(env);
        if (v == "8" || v == "16") l->tile = std::stoi(v);
    }
    // QJX_LB_FUSE=2: with 8 x 8 tiles, two RK4 steps per launch (halo 16).  At d = 101
    // it measured 14% slower than one step per launch: the recomputed halo work
    // (5.3x) costs more than the launch and the load of rho it saves
    if (const char *env = std::getenv("QJX_LB_FUSE"))
        if (std::string(env) == "2") l->fuse = 2;
    auto bail = [&](int code) {
        qjx_lindblad_destroy(l);
        return code;
    };
    std::vector<double> hb;
    build_bands(m, l->d, hb);
    const int d = l->d;
    const size_t n = nsq(l);
    if (hipStreamCreate(&l->stream) != hipSuccess) return bail(lb_fail(QJX_ERR_HIP, "hipStreamCreate failed"));
    if (hipMalloc((void **)&l->d_step, sizeof(long long)) != hipSuccess ||
        hipMalloc((void **)&l->bands, hb.size() * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&l->buf, 2 * n * sizeof(double2)) != hipSuccess ||
        hipMalloc((void **)&l->obs, ((size_t)4 + d) * sizeof(double)) != hipSuccess)
        return bail(lb_fail(QJX_ERR_NOMEM, "Lindblad buffers do not fit in device memory"));
    if (hipMemcpy(l->bands, hb.data(), hb.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return bail(lb_fail(QJX_ERR_HIP, "band upload failed"));
    double *bp = l->bands;
    l->B = qjx::LBands{bp, bp + d, bp + 2 * d, bp + 3 * d, bp + 4 * d, bp + 5 * d, bp + 6 * d, bp + 7 * d,
                       m->diss_gamma / double(d - 1), d};
    qjx::LDrive &D = l->D;
    D.step = l->d_step;
    D.T = l->T;
    D.h = l->h;
    D.E0 = m->prm_E;
    D.A = m->drv_ampl;
    D.freq = m->drv_freq;
    D.phase = drive->drv_phase;
    D.Eplus = m->prm_E + m->drv_ampl;
    D.Eminus = m->prm_E - m->drv_ampl;
    D.S = steps_per_period;
    D.type = drive->drv_type;
    D.per_stage = drive->per_stage;
    const long long zero = 0;
    if (hipMemcpy(l->d_step, &zero, sizeof zero, hipMemcpyHostToDevice) != hipSuccess)
        return bail(lb_fail(QJX_ERR_HIP, "counter upload failed"));
    // start in the pure state e_0 (start_type 0, start_state 0)
    std::vector<double2> r0(n, make_double2(0.0, 0.0));
    r0[0] = make_double2(1.0, 0.0);
    if (hipMemcpy(l->buf, r0.data(), n * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess)
        return bail(lb_fail(QJX_ERR_HIP, "rho upload failed"));
    *out = l;
    return QJX_OK;
}

extern "C" int qjx_lindblad_set_rho(qjx_lindblad *l, const double *rho)
{
    if (!l || !rho) return lb_fail(QJX_ERR_ARG, "null argument");
    LB_TRY(hipStreamSynchronize(l->stream));
    LB_TRY(hipMemcpy(rho_of(l), rho, nsq(l) * sizeof(double2), hipMemcpyHostToDevice));
    const long long zero = 0;      // a new initial state starts at t = 0
    LB_TRY(hipMemcpy(l->d_step, &zero, sizeof zero, hipMemcpyHostToDevice));
    return QJX_OK;
}

extern "C" int qjx_lindblad_get_rho(qjx_lindblad *l, double *rho)
{
    if (!l || !rho) return lb_fail(QJX_ERR_ARG, "null argument");
    LB_TRY(hipStreamSynchronize(l->stream));
    LB_TRY(hipMemcpy(rho, rho_of(l), nsq(l) * sizeof(double2), hipMemcpyDeviceToHost));
    return QJX_OK;
}

extern "C" int qjx_lindblad_run_periods(qjx_lindblad *l, int n_periods)
{
    if (!l) return lb_fail(QJX_ERR_ARG, "null argument");
    if (n_periods < 0) return lb_fail(QJX_ERR_ARG, "negative period count");
    if (n_periods == 0) return QJX_OK;
    // kChunk consecutive RK4 steps per graph launch (the step kernel reads the step
    // from the device counter), the remainder launched eagerly; every launch checked
    const long long total = (long long)n_periods * l->steps;
    long long done = 0;
    if (l->graphs && total >= kChunk) {
        if (!l->exec) {
            LB_TRY(hipGetLastError());             // an earlier launch error, as itself
            LB_TRY(hipStreamBeginCapture(l->stream, hipStreamCaptureModeThreadLocal));
            const int lr = launch_steps(l, kChunk, l->stream);   // an even number of launches: no copy
            const hipError_t le = lr == QJX_OK ? hipGetLastError() : hipErrorUnknown;
            hipGraph_t g = nullptr;
            const hipError_t ee = hipStreamEndCapture(l->stream, &g);   // always ends the capture
            if (le != hipSuccess || ee != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                (void)hipGetLastError();
                return lb_fail(QJX_ERR_HIP, std::string("capture of the RK4 steps failed: ") +
                                                hipGetErrorString(le != hipSuccess ? le : ee));
            }
            const hipError_t ie = hipGraphInstantiate(&l->exec, g, nullptr, nullptr, 0);
            if (ie != hipSuccess) {
                (void)hipGraphDestroy(g);
                l->exec = nullptr;
                return lb_fail(ie == hipErrorOutOfMemory ? QJX_ERR_NOMEM : QJX_ERR_HIP,
                               std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
            }
            l->graph = g;
        }
        for (; done + kChunk <= total; done += kChunk) {
            LB_TRY(hipGraphLaunch(l->exec, l->stream));
            // bound the launches in flight: wait for the previous period's work
            if ((done / kChunk) % 200 == 199) LB_TRY(hipStreamSynchronize(l->stream));
        }
    }
    for (; done < total; done += kChunk) {
        const int n = (int)std::min<long long>(kChunk, total - done);
        if (const int rc = launch_steps(l, n, l->stream)) return rc;
        LB_TRY(hipGetLastError());
    }
    LB_TRY(hipStreamSynchronize(l->stream));
    return QJX_OK;
}

// CQdiss's rho output (data.cpp:40-46 with save_sparse_complex_mtx, utils.cpp:266-310):
// every entry, one line each, "row col re im" with 1-based indices, setprecision(16)
// scientific.
extern "C" int qjx_lindblad_save_rho(qjx_lindblad *l, const char *path)
{
    if (!l || !path) return lb_fail(QJX_ERR_ARG, "null argument");
    std::vector<double2> r(nsq(l));
    LB_TRY(hipStreamSynchronize(l->stream));
    LB_TRY(hipMemcpy(r.data(), rho_of(l), r.size() * sizeof(double2), hipMemcpyDeviceToHost));
    std::ofstream o(path);
    if (!o.is_open()) return lb_fail(QJX_ERR_IO, std::string("unable to open file ") + path);
    o << std::setprecision(16) << std::scientific;
    const int d = l->d;
    for (int i = 0; i < d; i++)
        for (int j = 0; j < d; j++)
            o << i + 1 << " " << j + 1 << " " << r[(size_t)i * d + j].x << " " << r[(size_t)i * d + j].y << std::endl;
    if (!o.good()) return lb_fail(QJX_ERR_IO, std::string("unable to write file ") + path);
    return QJX_OK;
}

extern "C" int qjx_lindblad_observables(qjx_lindblad *l, double *mean, double *m2, double *energy,
                                        double *purity, double *adr)
{
    if (!l) return lb_fail(QJX_ERR_ARG, "null argument");
    const int d = l->d;
    qjx::lb_obs_kernel<<<1, 256, 0, l->stream>>>(l->B, l->Estat, rho_of(l), l->obs, l->obs + 4);
    LB_TRY(hipGetLastError());
    qjx::launch_purity(reinterpret_cast<const double *>(rho_of(l)), (size_t)d * d, l->obs + 3, l->stream);
    LB_TRY(hipGetLastError());
    std::vector<double> h((size_t)4 + d);
    LB_TRY(hipMemcpyAsync(h.data(), l->obs, h.size() * sizeof(double), hipMemcpyDeviceToHost, l->stream));
    LB_TRY(hipStreamSynchronize(l->stream));
    if (mean) *mean = h[0];
    if (m2) *m2 = h[1];
    if (energy) *energy = h[2];
    if (purity) *purity = h[3];
    if (adr)
        for (int k = 0; k < d; k++) adr[k] = h[4 + k];
    return QJX_OK;
}

extern "C" int qjx_lindblad_dim(const qjx_lindblad *l) { return l ? l->d : -1; }
