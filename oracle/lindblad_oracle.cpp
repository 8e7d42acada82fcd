/*
 * lindblad_oracle.cpp -- density-matrix (rho-space) RK4 integration of the dimer's
 * Lindblad equation, the statistical oracle for quantum-jump ensembles.
 * TEST INFRASTRUCTURE ONLY (see qjx_oracle.h).
 *
 *   d rho / dt = -i [H(t), rho] + g (L rho L^dag - 1/2 {L^dag L, rho}),  g = gamma / N,
 *   H(t) = H + E(t) H_drv,  E(t) = E + A on the first half period, E - A on the second,
 *
 * with H, H_drv, L the matrices of qjx_oracle.cpp (reference newdel.cpp:450-498,
 * 1454-1498).  The reference's Lindblad program (CQdiss_os_dimer/CQdiss/calcODE.cpp:
 * 75-118, 249-389) integrates the same equation in the SU(N) f-basis with RK4; a fixed
 * linear change of variables commutes with RK4, so the rho-space integration is the same
 * scheme.  Two drive rules:
 *   - CQdiss's (calcVectValue_real, calcODE.cpp:100-116): at every RK4 stage time
 *     t = p T + s h (+ h/2, + h), E(t) = E + A sin(omega t + phase) (dimer_drv_type 1)
 *     or E + A sign(sin(omega t + phase)) (type 0, a zero sine counting as positive);
 *     at each switch of the square drive one stage then takes the neighbouring half
 *     period's value (an O(h) error, as in the reference).  The sine is the Cody-Waite
 *     + Taylor code of the trajectory oracle (qjo_rk_sin), within 2 ulp of libm.  CQdiss
 *     fixes T = 2 pi whatever the frequency; here T = 2 pi / omega (equal at omega = 1).
 *   - qjl_evolve's: the square drive held constant over each step, steps aligned with
 *     the half periods, so the scheme is RK4-accurate.
 *
 * H and L are tridiagonal and L^dag L is pentadiagonal, so each right-hand side
 * costs O(d^2).
 */
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "qjx_oracle.h"

typedef std::complex<double> cplx;

namespace {

struct Band {
    int d;
    std::vector<double> hd, ho;      // H diagonal (with E H_drv added per call), off-diagonal
    std::vector<double> hdrv;        // H_drv diagonal
    std::vector<double> ld, lo;      // L: L_kk = ld[k], L_{k+1,k} = -lo[k], L_{k,k+1} = +lo[k]
    std::vector<double> q0, q1, q2;  // L^dag L bands: (k,k), (k,k+1), (k,k+2)
    double g;
};

void build_band(const qjo_model *m, Band *B)
{
    const int d = m->N + 1;
    B->d = d;
    const double U = m->prm_U / double(d - 1);
    B->hd.assign(d, 0.0);
    B->hdrv.assign(d, 0.0);
    B->ld.assign(d, 0.0);
    B->ho.assign(d, 0.0);
    B->lo.assign(d, 0.0);
    for (int k = 0; k < d; k++) {
        B->hd[k] = 2.0 * U * double(k * (k - 1) + (d - (k + 1)) * (d - (k + 1) - 1));
        B->hdrv[k] = double((d - (k + 1)) - k);
        B->ld[k] = double(k - (d - (k + 1)));
    }
    for (int k = 0; k + 1 < d; k++) {
        B->ho[k] = -m->prm_J * std::sqrt(double((d - (k + 1)) * (k + 1)));
        B->lo[k] = std::sqrt(double((d - (k + 1)) * (k + 1)));
    }
    auto Lel = [&](int i, int j) -> double {
        if (i == j) return B->ld[i];
        if (i == j + 1) return -B->lo[j];
        if (j == i + 1) return B->lo[i];
        return 0.0;
    };
    B->q0.assign(d, 0.0);
    B->q1.assign(d, 0.0);
    B->q2.assign(d, 0.0);
    for (int i = 0; i < d; i++)
        for (int off = 0; off <= 2 && i + off < d; off++) {
            const int j = i + off;
            double s = 0.0;
            for (int k = 0; k < d; k++) s += Lel(k, i) * Lel(k, j);
            (off == 0 ? B->q0 : off == 1 ? B->q1 : B->q2)[i] = s;
        }
    B->g = m->gamma / double(d - 1);
}

/* y = A x for a symmetric real band matrix given by up to 3 bands (x, y d x d, rows). */
inline cplx row_tri(const std::vector<double> &dg, const std::vector<double> &of, const cplx *col,
                    int i, int d, int stride)
{
    cplx s = dg[i] * col[(size_t)i * stride];
    if (i > 0) s += of[i - 1] * col[(size_t)(i - 1) * stride];
    if (i + 1 < d) s += of[i] * col[(size_t)(i + 1) * stride];
    return s;
}

/* rhs = -i[H, rho] + g (L rho L^T - 1/2 (Q rho + rho Q)), all real band matrices. */
void rhs(const Band &B, double E, const cplx *rho, cplx *out, std::vector<cplx> &tmp)
{
    const int d = B.d;
    std::vector<double> hd(d);
    for (int k = 0; k < d; k++) hd[k] = B.hd[k] + E * B.hdrv[k];
    /* L rho L^T: first T = L rho (rows), then out = T L^T (columns). */
    tmp.assign((size_t)d * d, 0.0);
    for (int i = 0; i < d; i++)
        for (int j = 0; j < d; j++) {
            cplx s = B.ld[i] * rho[(size_t)i * d + j];
            if (i > 0) s += -B.lo[i - 1] * rho[(size_t)(i - 1) * d + j];
            if (i + 1 < d) s += B.lo[i] * rho[(size_t)(i + 1) * d + j];
            tmp[(size_t)i * d + j] = s;
        }
    for (int i = 0; i < d; i++)
        for (int j = 0; j < d; j++) {
            /* (T L^T)_{ij} = sum_k T_ik L_jk */
            cplx lrl = B.ld[j] * tmp[(size_t)i * d + j];
            if (j > 0) lrl += -B.lo[j - 1] * tmp[(size_t)i * d + j - 1];
            if (j + 1 < d) lrl += B.lo[j] * tmp[(size_t)i * d + j + 1];
            /* (H rho)_{ij} and (rho H)_{ij} */
            cplx hr = hd[i] * rho[(size_t)i * d + j];
            if (i > 0) hr += B.ho[i - 1] * rho[(size_t)(i - 1) * d + j];
            if (i + 1 < d) hr += B.ho[i] * rho[(size_t)(i + 1) * d + j];
            cplx rh = rho[(size_t)i * d + j] * hd[j];
            if (j > 0) rh += rho[(size_t)i * d + j - 1] * B.ho[j - 1];
            if (j + 1 < d) rh += rho[(size_t)i * d + j + 1] * B.ho[j];
            /* (Q rho)_{ij}, (rho Q)_{ij}, Q pentadiagonal symmetric */
            cplx qr = B.q0[i] * rho[(size_t)i * d + j];
            if (i + 1 < d) qr += B.q1[i] * rho[(size_t)(i + 1) * d + j];
            if (i + 2 < d) qr += B.q2[i] * rho[(size_t)(i + 2) * d + j];
            if (i >= 1) qr += B.q1[i - 1] * rho[(size_t)(i - 1) * d + j];
            if (i >= 2) qr += B.q2[i - 2] * rho[(size_t)(i - 2) * d + j];
            cplx rq = rho[(size_t)i * d + j] * B.q0[j];
            if (j + 1 < d) rq += rho[(size_t)i * d + j + 1] * B.q1[j];
            if (j + 2 < d) rq += rho[(size_t)i * d + j + 2] * B.q2[j];
            if (j >= 1) rq += rho[(size_t)i * d + j - 1] * B.q1[j - 1];
            if (j >= 2) rq += rho[(size_t)i * d + j - 2] * B.q2[j - 2];
            out[(size_t)i * d + j] = cplx(0.0, -1.0) * (hr - rh) + B.g * (lrl - 0.5 * (qr + rq));
        }
}

}  // namespace

extern "C" {

/* Evolve rho (d x d interleaved complex, row-major) over n_periods periods with
 * steps_per_period RK4 steps, from period period0 on (its time is period0 T).
 * per_stage = 1: CQdiss's drive at every stage time (drv_type, drv_phase); 0: the
 * square drive constant over each step (each half period gets half of the steps). */
int qjl_evolve_drive(const qjo_model *m, int steps_per_period, int period0, int n_periods, int drv_type,
                     double drv_phase, int per_stage, double *rho_io)
{
    if (!m || steps_per_period < 2 || (steps_per_period % 2) || n_periods < 0) return -1;
    Band B;
    build_band(m, &B);
    const int d = B.d;
    const size_t n = (size_t)d * d;
    cplx *rho = reinterpret_cast<cplx *>(rho_io);
    std::vector<cplx> k1(n), k2(n), k3(n), k4(n), v(n), tmp;
    const double T = 2.0 * 3.1415926535897932384626433832795 / m->drv_freq;
    const double h = T / double(steps_per_period);
    auto drive = [&](int s, double t) {
        if (!per_stage) return s < steps_per_period / 2 ? m->prm_E + m->drv_ampl : m->prm_E - m->drv_ampl;
        const double sn = qjo_rk_sin(m->drv_freq * t + drv_phase);
        const double f = drv_type == 1 ? sn : (sn < 0.0 ? -1.0 : 1.0);
        return m->prm_E + m->drv_ampl * f;
    };
    for (int p = period0; p < period0 + n_periods; p++)
        for (int s = 0; s < steps_per_period; s++) {
            const double time = double(p) * T + double(s) * h;   /* calcODE.cpp:346 */
            rhs(B, drive(s, time), rho, k1.data(), tmp);
            for (size_t i = 0; i < n; i++) v[i] = rho[i] + 0.5 * h * k1[i];
            rhs(B, drive(s, time + h / 2.0), v.data(), k2.data(), tmp);
            for (size_t i = 0; i < n; i++) v[i] = rho[i] + 0.5 * h * k2[i];
            rhs(B, drive(s, time + h / 2.0), v.data(), k3.data(), tmp);
            for (size_t i = 0; i < n; i++) v[i] = rho[i] + h * k3[i];
            rhs(B, drive(s, time + h), v.data(), k4.data(), tmp);
            for (size_t i = 0; i < n; i++) rho[i] += (h / 6.0) * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]);
        }
    return 0;
}

int qjl_evolve(const qjo_model *m, int steps_per_period, int n_periods, double *rho_io)
{
    return qjl_evolve_drive(m, steps_per_period, 0, n_periods, 0, 0.0, 0, rho_io);
}

/* The Lindblad right-hand side itself (for pinning against a dense Liouvillian). */
int qjl_rhs(const qjo_model *m, double E, const double *rho_in, double *out)
{
    Band B;
    build_band(m, &B);
    std::vector<cplx> tmp;
    rhs(B, E, reinterpret_cast<const cplx *>(rho_in), reinterpret_cast<cplx *>(out), tmp);
    return 0;
}

}  // extern "C"
