/*
 * qjx_oracle.cpp -- FP64 CPU restatement of os_d QJX's QJ trajectory path for the
 * Bose-Hubbard dimer.  TEST INFRASTRUCTURE ONLY (see qjx_oracle.h): the checker for
 * the HIP engine and the CPU baseline timed by bench.py.  Never linked by the product.
 *
 * Each function cites the reference code it restates (paths relative to the
 * reference repository root, source/cpp/QJX/QJX/ unless stated otherwise).
 */
#include "qjx_oracle.h"

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef std::complex<double> cplx;

static const double kPi = 3.1415926535897932384626433832795;  /* config.h PI */

/* ------------------------------------------------------------------------- */
/* Dense complex matrix helpers (row-major n x n).                           */
/* ------------------------------------------------------------------------- */

static void matmul(int n, const cplx *A, const cplx *B, cplx *C)
{
    /* C = A B, i-k-j order; parallel over rows. */
#pragma omp parallel for schedule(static) if (n >= 128)
    for (int i = 0; i < n; i++) {
        cplx *c = C + (size_t)i * n;
        for (int j = 0; j < n; j++) c[j] = 0.0;
        for (int k = 0; k < n; k++) {
            const cplx a = A[(size_t)i * n + k];
            if (a == cplx(0.0, 0.0)) continue;
            const cplx *b = B + (size_t)k * n;
            for (int j = 0; j < n; j++) c[j] += a * b[j];
        }
    }
}

static double norm1(int n, const cplx *A)
{
    double best = 0.0;
    for (int j = 0; j < n; j++) {
        double s = 0.0;
        for (int i = 0; i < n; i++) s += std::abs(A[(size_t)i * n + j]);
        if (s > best) best = s;
    }
    return best;
}

/* Solve P X = Q in place (Q <- X) by LU with partial pivoting.  P is destroyed. */
static int lu_solve(int n, cplx *P, cplx *Q)
{
    std::vector<int> piv(n);
    for (int k = 0; k < n; k++) {
        int p = k;
        double best = std::abs(P[(size_t)k * n + k]);
        for (int i = k + 1; i < n; i++) {
            double v = std::abs(P[(size_t)i * n + k]);
            if (v > best) { best = v; p = i; }
        }
        if (best == 0.0) return -1;
        piv[k] = p;
        if (p != k) {
            for (int j = 0; j < n; j++) std::swap(P[(size_t)k * n + j], P[(size_t)p * n + j]);
            for (int j = 0; j < n; j++) std::swap(Q[(size_t)k * n + j], Q[(size_t)p * n + j]);
        }
        const cplx inv = 1.0 / P[(size_t)k * n + k];
#pragma omp parallel for schedule(static) if (n >= 128)
        for (int i = k + 1; i < n; i++) {
            cplx f = P[(size_t)i * n + k] * inv;
            if (f == cplx(0.0, 0.0)) continue;
            P[(size_t)i * n + k] = f;
            for (int j = k + 1; j < n; j++) P[(size_t)i * n + j] -= f * P[(size_t)k * n + j];
            for (int j = 0; j < n; j++) Q[(size_t)i * n + j] -= f * Q[(size_t)k * n + j];
        }
    }
    for (int k = n - 1; k >= 0; k--) {
        const cplx inv = 1.0 / P[(size_t)k * n + k];
        for (int j = 0; j < n; j++) Q[(size_t)k * n + j] *= inv;
#pragma omp parallel for schedule(static) if (n >= 128)
        for (int i = 0; i < k; i++) {
            cplx f = P[(size_t)i * n + k];
            if (f == cplx(0.0, 0.0)) continue;
            for (int j = 0; j < n; j++) Q[(size_t)i * n + j] -= f * Q[(size_t)k * n + j];
        }
    }
    return 0;
}

/* Matrix exponential by diagonal Pade approximants with scaling and squaring,
 * N. J. Higham, "The scaling and squaring method for the matrix exponential
 * revisited", SIAM J. Matrix Anal. Appl. 26 (2005) 1179-1193, Algorithm 2.3:
 * degree m in {3,5,7,9,13} picked from the 1-norm thresholds theta_m, and for
 * m = 13 the matrix is scaled by 2^-s so that ||A/2^s||_1 <= theta_13.
 * The reference calls Eigen's MatrixBase::exp(), which implements the same
 * published algorithm (qj_proc.cpp:55, 98). */
static int expm_pade(int n, const cplx *A, cplx *E)
{
    static const double b3[] = {120., 60., 12., 1.};
    static const double b5[] = {30240., 15120., 3360., 420., 30., 1.};
    static const double b7[] = {17297280., 8648640., 1995840., 277200., 25200., 1512., 56., 1.};
    static const double b9[] = {17643225600., 8821612800., 2075673600., 302702400., 30270240.,
                                2162160., 110880., 3960., 90., 1.};
    static const double b13[] = {64764752532480000., 32382376266240000., 7771770303897600.,
                                 1187353796428800., 129060195264000., 10559470521600.,
                                 670442572800., 33522128640., 1323241920., 40840800.,
                                 960960., 16380., 182., 1.};
    const size_t nn = (size_t)n * n;
    const double a1 = norm1(n, A);
    int m;
    const double *b;
    int s = 0;
    if (a1 <= 1.495585217958292e-2) { m = 3; b = b3; }
    else if (a1 <= 2.539398330063230e-1) { m = 5; b = b5; }
    else if (a1 <= 9.504178996162932e-1) { m = 7; b = b7; }
    else if (a1 <= 2.097847961257068e0) { m = 9; b = b9; }
    else {
        m = 13; b = b13;
        const double theta13 = 5.371920351148152e0;
        if (a1 > theta13) s = (int)std::ceil(std::log2(a1 / theta13));
        if (s < 0) s = 0;
    }
    std::vector<cplx> X(A, A + nn);
    if (s > 0) {
        const double f = std::ldexp(1.0, -s);
        for (size_t i = 0; i < nn; i++) X[i] *= f;
    }
    std::vector<cplx> X2(nn), U(nn), V(nn), T(nn);
    matmul(n, X.data(), X.data(), X2.data());
    if (m < 13) {
        /* U = X * sum_{odd j} b_j X^{j-1}, V = sum_{even j} b_j X^j, powers of X2. */
        std::vector<cplx> P(nn), Uo(nn, 0.0);
        for (size_t i = 0; i < nn; i++) V[i] = 0.0;
        /* P = I */
        for (size_t i = 0; i < nn; i++) P[i] = 0.0;
        for (int i = 0; i < n; i++) P[(size_t)i * n + i] = 1.0;
        for (int j = 0; 2 * j <= m; j++) {
            for (size_t i = 0; i < nn; i++) {
                V[i] += b[2 * j] * P[i];
                if (2 * j + 1 <= m) Uo[i] += b[2 * j + 1] * P[i];
            }
            if (2 * (j + 1) <= m) {
                matmul(n, P.data(), X2.data(), T.data());
                P.swap(T);
            }
        }
        matmul(n, X.data(), Uo.data(), U.data());
    } else {
        std::vector<cplx> X4(nn), X6(nn), W(nn);
        matmul(n, X2.data(), X2.data(), X4.data());
        matmul(n, X4.data(), X2.data(), X6.data());
        /* W = X6 (b13 X6 + b11 X4 + b9 X2) + b7 X6 + b5 X4 + b3 X2 + b1 I */
        for (size_t i = 0; i < nn; i++) T[i] = b[13] * X6[i] + b[11] * X4[i] + b[9] * X2[i];
        matmul(n, X6.data(), T.data(), W.data());
        for (size_t i = 0; i < nn; i++) W[i] += b[7] * X6[i] + b[5] * X4[i] + b[3] * X2[i];
        for (int i = 0; i < n; i++) W[(size_t)i * n + i] += b[1];
        matmul(n, X.data(), W.data(), U.data());
        /* V = X6 (b12 X6 + b10 X4 + b8 X2) + b6 X6 + b4 X4 + b2 X2 + b0 I */
        for (size_t i = 0; i < nn; i++) T[i] = b[12] * X6[i] + b[10] * X4[i] + b[8] * X2[i];
        matmul(n, X6.data(), T.data(), V.data());
        for (size_t i = 0; i < nn; i++) V[i] += b[6] * X6[i] + b[4] * X4[i] + b[2] * X2[i];
        for (int i = 0; i < n; i++) V[(size_t)i * n + i] += b[0];
    }
    /* (V - U) E = (V + U) */
    std::vector<cplx> P(nn);
    for (size_t i = 0; i < nn; i++) {
        P[i] = V[i] - U[i];
        E[i] = V[i] + U[i];
    }
    if (lu_solve(n, P.data(), E) != 0) return -1;
    for (int k = 0; k < s; k++) {
        matmul(n, E, E, T.data());
        std::memcpy(E, T.data(), nn * sizeof(cplx));
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Model (newdel.cpp:3-37, 450-498, 1454-1498, 2285-2388).                   */
/* ------------------------------------------------------------------------- */

struct Model {
    int N = 0, d = 0;
    double T = 0.0;
    std::vector<cplx> H, Hdrv, L, Hqj[2];
    std::vector<cplx> Hp;                 /* H_part = H - i/2 gamma' L^dag L      */
};

static void build_model(const qjo_model *p, Model *m)
{
    const int N = p->N, d = N + 1;
    m->N = N;
    m->d = d;
    m->T = 2.0 * kPi / p->drv_freq;                       /* newdel.cpp:31 */
    const size_t dd = (size_t)d * d;
    m->H.assign(dd, 0.0);
    m->Hdrv.assign(dd, 0.0);
    m->L.assign(dd, 0.0);
    const double U = p->prm_U / double(d - 1);            /* newdel.cpp:459 */
    for (int k = 0; k < d; k++) {                         /* newdel.cpp:480-486 */
        m->H[(size_t)k * d + k] = 2.0 * U * double(k * (k - 1) + (d - (k + 1)) * (d - (k + 1) - 1));
        m->Hdrv[(size_t)k * d + k] = double((d - (k + 1)) - k);
        m->L[(size_t)k * d + k] = double(k - (d - (k + 1))); /* newdel.cpp:1482-1488 */
    }
    for (int k = 0; k < d - 1; k++) {                     /* newdel.cpp:488-495, 1490-1497 */
        const double h = p->prm_J * std::sqrt(double((d - (k + 1)) * (k + 1)));
        m->H[(size_t)(k + 1) * d + k] -= h;
        m->H[(size_t)k * d + k + 1] -= h;
        const double l = std::sqrt(double((d - (k + 1)) * (k + 1)));
        m->L[(size_t)(k + 1) * d + k] -= l;
        m->L[(size_t)k * d + k + 1] += l;
    }
    /* diss_part = gamma' L^dag L (newdel.cpp:2327-2342); H_part = H - i/2 diss_part;
     * Hqj[b] = -i (H_part + E_b H_drv) (newdel.cpp:2346-2383). */
    const double g = p->gamma / double(d - 1);
    std::vector<cplx> LdL(dd, 0.0);
    for (int i = 0; i < d; i++)
        for (int j = 0; j < d; j++) {
            cplx s = 0.0;
            for (int k = 0; k < d; k++) s += std::conj(m->L[(size_t)k * d + i]) * m->L[(size_t)k * d + j];
            LdL[(size_t)i * d + j] = g * s;
        }
    const double E[2] = {p->prm_E + p->drv_ampl, p->prm_E - p->drv_ampl};
    m->Hp.assign(dd, 0.0);
    for (size_t idx = 0; idx < dd; idx++)
        m->Hp[idx] = cplx(m->H[idx].real() + 0.5 * LdL[idx].imag(), 0.0 - 0.5 * LdL[idx].real());
    for (int b = 0; b < 2; b++) {
        m->Hqj[b].assign(dd, 0.0);
        for (size_t idx = 0; idx < dd; idx++) {
            const double hre = m->H[idx].real() + 0.5 * LdL[idx].imag();
            const double him = 0.0 - 0.5 * LdL[idx].real();
            m->Hqj[b][idx] = cplx(him + E[b] * 0.0, -(hre + E[b] * m->Hdrv[idx].real()));
        }
    }
}

/* ------------------------------------------------------------------------- */
/* RNG: MKL VSL MCG59, a = 13^13, m = 2^59, seed 777, leap-frog(k, mns)       */
/* (newdel_exp.cpp:244-296).  Output n of stream k is x_{k + n mns} / 2^59 with */
/* x_j = 777 a^j mod 2^59; the conversion is the correctly rounded double of x */
/* times 2^-59 (pinned by MKL golden vectors).                                 */
/* ------------------------------------------------------------------------- */

static const uint64_t kMcgA = 302875106592253ULL;          /* 13^13 */
static const uint64_t kMcgMask = (1ULL << 59) - 1ULL;

static uint64_t mcg_pow(uint64_t a, uint64_t e)
{
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = (r * a) & kMcgMask;
        a = (a * a) & kMcgMask;
        e >>= 1;
    }
    return r;
}

struct Rng {
    uint64_t x, stride;
    double next()
    {
        double u = (double)x * 0x1p-59;
        x = (x * stride) & kMcgMask;
        return u;
    }
};

static Rng rng_stream(int k, int mns)
{
    Rng r;
    r.x = (777ULL * mcg_pow(kMcgA, (uint64_t)k)) & kMcgMask;
    r.stride = mcg_pow(kMcgA, (uint64_t)mns);
    return r;
}

extern "C" void qjo_mcg59_uniform(int k, int mns, int n, double *out)
{
    Rng r = rng_stream(k, mns);
    for (int i = 0; i < n; i++) out[i] = r.next();
}

/* ------------------------------------------------------------------------- */
/* Ensemble.                                                                  */
/* ------------------------------------------------------------------------- */

struct qjo_ens {
    qjo_model p;
    Model m;
    int M = 0, seed = 0, mns = 0, start_type = 0, start_state = 0, jump = 0, threads = 1;
    std::vector<double> dt;                 /* dt[i] of level i                    */
    std::vector<cplx> U[2];                 /* U[b][lvl * d*d ...]                  */
    std::vector<cplx> psi, psi_aux;         /* phi_all, phi_all_aux (M x d)         */
    std::vector<double> t, eta, mean_start;
    std::vector<int> njumps;
    std::vector<int> pos;                   /* top-level steps completed            */
    std::vector<Rng> rng;
    std::vector<std::vector<double>> jt, jn, je;
    std::vector<std::vector<int>> jtype;
    std::vector<int64_t> matvecs, jumps_done;
    /* Runge-Kutta propagation (prop_id 1), set by qjo_rk_setup */
    int prop = 0, rk_ns = 0, drv_type = 0;
    double drv_phase = 0.0, rk_step = 0.0;
    std::vector<double> b_dre, b_dim, b_o1re, b_o1im, b_o2im, b_drv, b_Ld, b_Lm, b_Lp;
    struct Lpn *lpn = nullptr;              /* LPN bookkeeping, set by qjo_lpn_setup */
    ~qjo_ens();
    std::string err;
    volatile int failed = 0;
};

static void set_err(qjo_ens *e, const std::string &s)
{
#pragma omp critical(qjo_err)
    {
        if (!e->failed) e->err = s;
        e->failed = 1;
    }
}

extern "C" int qjo_model_matrices(const qjo_model *p, double *H, double *Hdrv, double *L,
                                  double *Hqj0, double *Hqj1)
{
    Model m;
    build_model(p, &m);
    const size_t bytes = (size_t)m.d * m.d * sizeof(cplx);
    if (H) std::memcpy(H, m.H.data(), bytes);
    if (Hdrv) std::memcpy(Hdrv, m.Hdrv.data(), bytes);
    if (L) std::memcpy(L, m.L.data(), bytes);
    if (Hqj0) std::memcpy(Hqj0, m.Hqj[0].data(), bytes);
    if (Hqj1) std::memcpy(Hqj1, m.Hqj[1].data(), bytes);
    return 0;
}

extern "C" int qjo_expm(int n, const double *A, double *E)
{
    return expm_pade(n, reinterpret_cast<const cplx *>(A), reinterpret_cast<cplx *>(E));
}

/* Split tree (core.cpp:9880-9926, qj_proc.cpp:4-120), generalised top level:
 * level 0 has S/2 steps of dt_0 = T/S, level i > 0 has 2 steps of dt_{i-1}/2.
 * With S = 4 this is exactly QJX (dt_i = T / 2^(i+2)). */
static void build_tree(qjo_ens *e, int build)
{
    const int d = e->m.d, D = e->p.D;
    e->dt.resize(D);
    double dt = e->m.T / double(e->p.S);
    for (int i = 0; i < D; i++) {
        e->dt[i] = dt;
        dt *= 0.5;
    }
    const size_t dd = (size_t)d * d;
    for (int b = 0; b < 2; b++) {
        e->U[b].assign(dd * D, 0.0);
        if (!build) continue;
        std::vector<cplx> A(dd);
        for (int i = 0; i < D; i++) {
            for (size_t idx = 0; idx < dd; idx++)
                A[idx] = cplx(e->m.Hqj[b][idx].real() * e->dt[i], e->m.Hqj[b][idx].imag() * e->dt[i]);
            expm_pade(d, A.data(), e->U[b].data() + dd * i);
        }
    }
}

extern "C" qjo_ens *qjo_create(const qjo_model *p, int num_traj, int seed, int mns,
                               int start_type, int start_state, int jump, int num_threads,
                               int build_propagators)
{
    if (!p || p->N < 1 || num_traj < 1 || p->S < 2 || (p->S % 2) || p->D < 1) return nullptr;
    qjo_ens *e = new qjo_ens;
    e->p = *p;
    build_model(p, &e->m);
    e->M = num_traj;
    e->seed = seed;
    e->mns = mns;
    e->start_type = start_type;
    e->start_state = start_state;
    e->jump = jump;
    e->threads = num_threads > 0 ? num_threads : 1;
#ifdef _OPENMP
    omp_set_num_threads(e->threads);
#endif
    build_tree(e, build_propagators);
    const int d = e->m.d, M = num_traj;
    /* init_basic_data (newdel_exp.cpp:332-375), init_start_state (656-679) */
    e->psi.assign((size_t)M * d, 0.0);
    e->psi_aux.assign((size_t)M * d, 0.0);
    e->t.assign(M, 0.0);
    e->eta.assign(M, 0.0);
    e->mean_start.assign(M, 0.0);
    e->njumps.assign(M, 0);
    e->pos.assign(M, 0);
    e->jt.assign(M, {});
    e->jn.assign(M, {});
    e->je.assign(M, {});
    e->jtype.assign(M, {});
    e->matvecs.assign(M, 0);
    e->jumps_done.assign(M, 0);
    e->rng.resize(M);
    for (int tr = 0; tr < M; tr++) {
        e->rng[tr] = rng_stream(seed + tr, mns);
        cplx *ph = &e->psi[(size_t)tr * d];
        if (start_type == 0) ph[start_state] = 1.0;
        else for (int i = 0; i < d; i++) ph[i] = std::sqrt(1.0 / double(d));
    }
    return e;
}

extern "C" void qjo_destroy(qjo_ens *e) { delete e; }
extern "C" const char *qjo_error(const qjo_ens *e) { return e ? e->err.c_str() : "null handle"; }

extern "C" int qjo_set_propagator(qjo_ens *e, int b, int lvl, const double *U)
{
    if (b < 0 || b > 1 || lvl < 0 || lvl >= e->p.D) return -1;
    const size_t dd = (size_t)e->m.d * e->m.d;
    std::memcpy(e->U[b].data() + dd * lvl, U, dd * sizeof(cplx));
    return 0;
}

extern "C" int qjo_get_propagator(const qjo_ens *e, int b, int lvl, double *U)
{
    if (b < 0 || b > 1 || lvl < 0 || lvl >= e->p.D) return -1;
    const size_t dd = (size_t)e->m.d * e->m.d;
    std::memcpy(U, e->U[b].data() + dd * lvl, dd * sizeof(cplx));
    return 0;
}

/* norm_square (experiment.cpp:1263-1271) */
static double norm_square(const cplx *x, int d)
{
    double s = 0.0;
    for (int i = 0; i < d; i++) s += x[i].real() * x[i].real() + x[i].imag() * x[i].imag();
    return s;
}

/* prop_step: row-major zgemv, res = U x (qj_proc.cpp:264-269) */
static void prop_step(const cplx *x, const cplx *U, cplx *res, int d)
{
    for (int i = 0; i < d; i++) {
        const cplx *row = U + (size_t)i * d;
        double re = 0.0, im = 0.0;
        for (int j = 0; j < d; j++) {
            re += row[j].real() * x[j].real() - row[j].imag() * x[j].imag();
            im += row[j].real() * x[j].imag() + row[j].imag() * x[j].real();
        }
        res[i] = cplx(re, im);
    }
}

/* recovery (experiment.cpp:1273-1382), one dissipator (num_diss = 1, g = 1). */
static void recovery(qjo_ens *e, int tr, int is_obs)
{
    const int d = e->m.d;
    cplx *phi = &e->psi_aux[(size_t)tr * d];
    double nrm = std::sqrt(norm_square(phi, d));
    for (int i = 0; i < d; i++) phi[i] = cplx(phi[i].real() / nrm, phi[i].imag() / nrm);
    const double ran = e->rng[tr].next();
    std::vector<cplx> res(d);
    prop_step(phi, e->m.L.data(), res.data(), d);         /* L_i phi            */
    double gnorm = norm_square(res.data(), d) * 1.0;       /* g_i = 1            */
    double tmp = gnorm;
    double r = ran * tmp;
    int index = 0;
    (void)r;                                                /* k = 1: index 0     */
    const double sc = std::sqrt(gnorm / 1.0);
    for (int i = 0; i < d; i++) phi[i] = cplx(res[i].real() / sc, res[i].imag() / sc);
    if (e->jump > 0 && is_obs == 1) e->jtype[tr].push_back(index);
    e->jumps_done[tr] += 1;
}

/* one_period_branch (qj_proc.cpp:271-343) as an explicit recursion over levels. */
static void level_run(qjo_ens *e, int tr, int b, int lvl, int steps, int is_obs)
{
    const int d = e->m.d, D = e->p.D;
    const size_t dd = (size_t)d * d;
    cplx *phi = &e->psi[(size_t)tr * d];
    cplx *aux = &e->psi_aux[(size_t)tr * d];
    const cplx *U = e->U[b].data() + dd * lvl;
    const double dt = e->dt[lvl];
    double &t = e->t[tr];
    double &eta = e->eta[tr];
    for (int c = 0; c < steps; c++) {
        if (e->failed) return;
        prop_step(phi, U, aux, d);
        e->matvecs[tr] += 1;
        const double n = norm_square(aux, d);
        if (n > 1.0) {                                      /* experiment.cpp:1247-1251 */
            char buf[160];
            std::snprintf(buf, sizeof buf, "norm more than 1 (trajectory %d, norm %.17g)", tr, n);
            set_err(e, buf);
            return;
        }
        const bool crossed = !(n > eta);                   /* experiment.cpp:1253-1260 */
        if (lvl == D - 1) {
            if (crossed) {
                if (e->jump > 0 && is_obs == 1) {
                    e->jt[tr].push_back(t + dt);
                    e->jn[tr].push_back(n);
                    e->je[tr].push_back(eta);
                    e->njumps[tr]++;
                }
                recovery(e, tr, is_obs);
                eta = e->rng[tr].next();
                while (eta == 0.0) eta = e->rng[tr].next();
            }
            std::memcpy(phi, aux, sizeof(cplx) * d);
            t += dt;
        } else {
            if (crossed) {
                level_run(e, tr, b, lvl + 1, 2, is_obs);
                t -= dt;
            } else {
                std::memcpy(phi, aux, sizeof(cplx) * d);
            }
            t += dt;
        }
    }
}

/* ex_period (core.cpp:84-95): both branches, top level S/2 steps each. */
static void one_period(qjo_ens *e, int tr, int is_obs)
{
    for (int b = 0; b < 2; b++) level_run(e, tr, b, 0, e->p.S / 2, is_obs);
    e->pos[tr] += e->p.S;
}

/* n top-level steps of one trajectory from its current position (the same
 * sequence as whole periods, cut at top-level step boundaries). */
static void top_steps(qjo_ens *e, int tr, int n, int is_obs)
{
    const int half = e->p.S / 2;
    for (int k = 0; k < n && !e->failed; k++) {
        const int b = (e->pos[tr] / half) % 2;
        level_run(e, tr, b, 0, 1, is_obs);
        e->pos[tr] += 1;
    }
}

/* ------------------------------------------------------------------------- */
/* Runge-Kutta trajectories (prop_id 1): rk_period (core.cpp:439-450),       */
/* rk_step_dimer (core.cpp:11886-11948), rk_int_dimer (core.cpp:11592-11619), */
/* rk_right_part_dimer (core.cpp:11033-11115), set_init_args / arg_upd /     */
/* rk_final / rk_recovery / save_phi_prev / restore_from_prev                */
/* (rk_proc.cpp:124-345), init_rk (rk_proc.cpp:56-63).                        */
/*                                                                           */
/* The reference leaves the floating-point order to MKL's zgemv and to the   */
/* compiler; this restatement fixes one (rounding-level differences only):    */
/*  - each multiply-add is one fma; the products H_part x and H_drv x run over */
/*    the band |i - j| <= 2 in ascending j (every other entry of H_part and    */
/*    H_drv is an exact zero, so the dense zgemv adds the same nonzero terms;  */
/*    qjo_rk_setup checks that);                                               */
/*  - the RK4 update is phi + acc * (step / 6) with one fma;                   */
/*  - ||x||^2 sums 64 chunks of E consecutive entries (E = the smallest power  */
/*    of two >= 2 with 64 E >= d), each in ascending order with fma, and adds  */
/*    the 64 partial sums by a butterfly over lane distances 32, 16, .., 1;     */
/*  - sin (dimer_drv_type 1) is rk_sin below, not the C library's.            */
/* Reproduced as the reference has them: the stage arguments use the global   */
/* rk_step whatever step rk_int is called with (arg_upd reads ed->rk_step,    */
/* rk_proc.cpp:150-173), time advances by step * 0.5 twice per integration,  */
/* a crossing restores the step-start state, moves time back by the step,     */
/* integrates the linear-interpolation fraction, jumps, draws ran then eta    */
/* (redrawn while 0) and integrates the rest (recursively: a loop here).      */
/* ------------------------------------------------------------------------- */

/* The GPU kernel's layout fixes the summation tree of ||x||^2: P lanes per
 * trajectory (32 for 64 < d <= 128, else 64), lane l holding entries
 * [l E, l E + E) with E the smallest power of two such that P E >= d; each lane sums
 * its entries with fma, then the butterfly adds partners at lane distances P/2..1. */
static int rk_group(int d) { return d > 64 && d <= 128 ? 32 : 64; }

static int rk_lanes(int d)
{
    const int P = rk_group(d);
    int E = 1;
    while (P * E < d) E *= 2;
    return E;
}

static double rk_norm(const cplx *x, int d)
{
    const int P = rk_group(d), E = rk_lanes(d);
    double part[64];
    for (int l = 0; l < P; l++) {
        double p = 0.0;
        for (int j = 0; j < E && l * E + j < d; j++) {
            const cplx &v = x[l * E + j];
            p = std::fma(v.real(), v.real(), p);
            p = std::fma(v.imag(), v.imag(), p);
        }
        part[l] = p;
    }
    for (int off = P / 2; off > 0; off >>= 1) {
        double nxt[64];
        for (int l = 0; l < P; l++) nxt[l] = part[l] + part[l ^ off];
        std::memcpy(part, nxt, sizeof(double) * P);
    }
    return part[0];
}

/* sin x: quadrant k = nearest integer to x 2/pi, r = x - k pi/2 with pi/2 split
 * into 33 + 33 + 53 significant bits (Cody & Waite, "Software Manual for the
 * Elementary Functions", 1980; k pi/2 parts are exact for |k| < 2^20), then the
 * Taylor series of sin r or cos r on |r| <= pi/4 (truncation < 1e-26) by Horner. */
static double rk_sin(double x)
{
    const double k = std::nearbyint(x * 0x1.45f306dc9c883p-1);
    double r = x - k * 0x1.921fb54400000p+0;
    r = r - k * 0x1.0b4611a600000p-34;
    r = r - k * 0x1.3198a2e037073p-69;
    const double z = r * r;
    double ps = -0x1.761b41316381ap-75;
    ps = std::fma(ps, z, 0x1.71b8ef6dcf572p-66);
    ps = std::fma(ps, z, -0x1.2f49b46814157p-57);
    ps = std::fma(ps, z, 0x1.952c77030ad4ap-49);
    ps = std::fma(ps, z, -0x1.ae7f3e733b81fp-41);
    ps = std::fma(ps, z, 0x1.6124613a86d09p-33);
    ps = std::fma(ps, z, -0x1.ae64567f544e4p-26);
    ps = std::fma(ps, z, 0x1.71de3a556c734p-19);
    ps = std::fma(ps, z, -0x1.a01a01a01a01ap-13);
    ps = std::fma(ps, z, 0x1.1111111111111p-7);
    ps = std::fma(ps, z, -0x1.5555555555555p-3);
    double pc = 0x1.f2cf01972f578p-80;
    pc = std::fma(pc, z, -0x1.0ce396db7f853p-70);
    pc = std::fma(pc, z, 0x1.e542ba4020225p-62);
    pc = std::fma(pc, z, -0x1.6827863b97d97p-53);
    pc = std::fma(pc, z, 0x1.ae7f3e733b81fp-45);
    pc = std::fma(pc, z, -0x1.93974a8c07c9dp-37);
    pc = std::fma(pc, z, 0x1.1eed8eff8d898p-29);
    pc = std::fma(pc, z, -0x1.27e4fb7789f5cp-22);
    pc = std::fma(pc, z, 0x1.a01a01a01a01ap-16);
    pc = std::fma(pc, z, -0x1.6c16c16c16c17p-10);
    pc = std::fma(pc, z, 0x1.5555555555555p-5);
    pc = std::fma(pc, z, -0x1.0000000000000p-1);
    const double s = std::fma(r * z, ps, r);
    const double c = std::fma(z, pc, 1.0);
    const int q = (int)((long long)k & 3);
    return q == 0 ? s : q == 1 ? c : q == 2 ? -s : -c;
}

/* E(t) of rk_right_part_dimer (core.cpp:11061-11088): any type but 1 is the
 * square drive. */
static double rk_drive(const qjo_ens *e, double t)
{
    if (e->drv_type == 1) return e->p.prm_E + e->p.drv_ampl * rk_sin(e->p.drv_freq * t + e->drv_phase);
    const double mt = std::fmod(t, e->m.T);
    return mt < e->m.T * 0.5 ? e->p.prm_E + e->p.drv_ampl : e->p.prm_E - e->p.drv_ampl;
}

/* k = -i (H_part + E H_drv) x: k.re = +(nd.im + E drv.im), k.im = -(nd.re + E drv.re)
 * with nd = H_part x, drv = H_drv x (core.cpp:11090-11114). */
static void rk_rhs(const qjo_ens *e, double E, const cplx *x, cplx *k)
{
    const int d = e->m.d;
    for (int i = 0; i < d; i++) {
        double re = 0.0, im = 0.0;
        auto mac = [&](double a, double b, const cplx &v) {
            re = std::fma(a, v.real(), re);
            re = std::fma(-b, v.imag(), re);
            im = std::fma(a, v.imag(), im);
            im = std::fma(b, v.real(), im);
        };
        auto mac_im = [&](double b, const cplx &v) {      /* entry 0 + i b */
            re = std::fma(-b, v.imag(), re);
            im = std::fma(b, v.real(), im);
        };
        if (i >= 2) mac_im(e->b_o2im[i - 2], x[i - 2]);
        if (i >= 1) mac(e->b_o1re[i - 1], e->b_o1im[i - 1], x[i - 1]);
        mac(e->b_dre[i], e->b_dim[i], x[i]);
        if (i + 1 < d) mac(e->b_o1re[i], e->b_o1im[i], x[i + 1]);
        if (i + 2 < d) mac_im(e->b_o2im[i], x[i + 2]);
        const double dre = e->b_drv[i] * x[i].real();
        const double dim = e->b_drv[i] * x[i].imag();
        k[i] = cplx(std::fma(E, dim, im), -std::fma(E, dre, re));
    }
}

struct RkWork {
    std::vector<cplx> nw, arg, k, acc;
    explicit RkWork(int d) : nw(d), arg(d), k(d), acc(d) {}
};

/* rk_int_dimer: out = RK4(phi, step) from time t; t advances by step * 0.5 twice. */
static void rk_int(const qjo_ens *e, double &t, double step, const cplx *phi, cplx *out, RkWork &w)
{
    const int d = e->m.d;
    const double h = 0.5 * e->rk_step, H = 1.0 * e->rk_step;
    cplx *arg = w.arg.data(), *k = w.k.data(), *acc = w.acc.data();
    rk_rhs(e, rk_drive(e, t), phi, k);
    for (int i = 0; i < d; i++) {
        acc[i] = k[i];
        arg[i] = cplx(std::fma(h, k[i].real(), phi[i].real()), std::fma(h, k[i].imag(), phi[i].imag()));
    }
    t = t + step * 0.5;
    rk_rhs(e, rk_drive(e, t), arg, k);
    for (int i = 0; i < d; i++) {
        acc[i] = cplx(std::fma(2.0, k[i].real(), acc[i].real()), std::fma(2.0, k[i].imag(), acc[i].imag()));
        arg[i] = cplx(std::fma(h, k[i].real(), phi[i].real()), std::fma(h, k[i].imag(), phi[i].imag()));
    }
    rk_rhs(e, rk_drive(e, t), arg, k);
    for (int i = 0; i < d; i++) {
        acc[i] = cplx(std::fma(2.0, k[i].real(), acc[i].real()), std::fma(2.0, k[i].imag(), acc[i].imag()));
        arg[i] = cplx(std::fma(H, k[i].real(), phi[i].real()), std::fma(H, k[i].imag(), phi[i].imag()));
    }
    t = t + step * 0.5;
    rk_rhs(e, rk_drive(e, t), arg, k);
    const double c6 = step / 6.0;
    for (int i = 0; i < d; i++) {
        const double ar = acc[i].real() + k[i].real(), ai = acc[i].imag() + k[i].imag();
        out[i] = cplx(std::fma(ar, c6, phi[i].real()), std::fma(ai, c6, phi[i].imag()));
    }
}

/* rk_recovery (rk_proc.cpp:198-306) with one dissipator: normalise, draw ran,
 * phi = L phi / ||L phi||. */
static void rk_recovery(qjo_ens *e, int tr, cplx *phi, double nsq, int is_obs, RkWork &w)
{
    const int d = e->m.d;
    const double nrm = std::sqrt(nsq);
    for (int i = 0; i < d; i++) phi[i] = cplx(phi[i].real() / nrm, phi[i].imag() / nrm);
    (void)e->rng[tr].next();                                /* ran; index 0 */
    cplx *res = w.k.data();
    for (int i = 0; i < d; i++) {
        double re = 0.0, im = 0.0;
        if (i >= 1) {
            re = std::fma(e->b_Lm[i], phi[i - 1].real(), re);
            im = std::fma(e->b_Lm[i], phi[i - 1].imag(), im);
        }
        re = std::fma(e->b_Ld[i], phi[i].real(), re);
        im = std::fma(e->b_Ld[i], phi[i].imag(), im);
        if (i + 1 < d) {
            re = std::fma(e->b_Lp[i], phi[i + 1].real(), re);
            im = std::fma(e->b_Lp[i], phi[i + 1].imag(), im);
        }
        res[i] = cplx(re, im);
    }
    const double g = rk_norm(res, d) * 1.0;
    const double sg = std::sqrt(g / 1.0);
    for (int i = 0; i < d; i++) phi[i] = cplx(res[i].real() / sg, res[i].imag() / sg);
    if (e->jump > 0 && is_obs == 1) e->jtype[tr].push_back(0);
    e->jumps_done[tr] += 1;
}

/* rk_step_dimer as a loop over its tail recursion.  `cached` carries ||phi||^2
 * of the current state between steps (recomputing it gives the same bits). */
static void rk_step(qjo_ens *e, int tr, double step, int is_obs, double &cached, bool &have, RkWork &w)
{
    const int d = e->m.d;
    cplx *phi = &e->psi[(size_t)tr * d];
    double &t = e->t[tr];
    double &eta = e->eta[tr];
    for (int crossings = 0;; crossings++) {
        const double prev = have ? cached : rk_norm(phi, d);
        rk_int(e, t, step, phi, w.nw.data(), w);
        e->matvecs[tr] += 4;
        const double n = rk_norm(w.nw.data(), d);
        if (n > 1.0) {                                      /* experiment.cpp:1247-1251 */
            char buf[160];
            std::snprintf(buf, sizeof buf, "norm more than 1 (trajectory %d, norm %.17g)", tr, n);
            set_err(e, buf);
            return;
        }
        /* the reference recurses without end on a NaN norm; stop instead */
        if (!(n == n) || crossings >= (1 << 16)) {
            set_err(e, "non-finite norm or endless crossings in one RK step");
            return;
        }
        if (n > eta) {                                      /* is_norm_crossed == 0 */
            std::memcpy(phi, w.nw.data(), sizeof(cplx) * d);
            cached = n;
            have = true;
            return;
        }
        const double diff = prev - n;
        const double bp = (prev - eta) / diff;
        const double ep = 1.0 - bp;
        const double bstep = step * bp, estep = step * ep;
        t = t - step;                                       /* restore_from_prev */
        rk_int(e, t, bstep, phi, w.nw.data(), w);
        e->matvecs[tr] += 4;
        std::memcpy(phi, w.nw.data(), sizeof(cplx) * d);
        const double jn = rk_norm(phi, d);
        if (e->jump > 0 && is_obs == 1) {
            e->jt[tr].push_back(t);
            e->jn[tr].push_back(jn);
            e->je[tr].push_back(eta);
            e->njumps[tr]++;
        }
        rk_recovery(e, tr, phi, jn, is_obs, w);
        eta = e->rng[tr].next();
        while (eta == 0.0) eta = e->rng[tr].next();
        have = false;
        step = estep;
    }
}

/* n RK steps of one trajectory; step s of the run has period_id = pos / rk_ns,
 * step_id = pos % rk_ns and starts at period_id T + step_id rk_step. */
static void rk_steps(qjo_ens *e, int tr, int n, int is_obs)
{
    RkWork w(e->m.d);
    double cached = 0.0;
    bool have = false;
    for (int s = 0; s < n && !e->failed; s++) {
        const int pos = e->pos[tr];
        const int period_id = pos / e->rk_ns, step_id = pos % e->rk_ns;
        e->t[tr] = double(period_id) * e->m.T + double(step_id) * e->rk_step;
        rk_step(e, tr, e->rk_step, is_obs, cached, have, w);
        e->pos[tr] += 1;
    }
}

extern "C" int qjo_rk_setup(qjo_ens *e, int rk_ns, int drv_type, double drv_phase)
{
    if (!e || rk_ns < 1) return -1;
    const int d = e->m.d;
    const Model &m = e->m;
    auto hp = [&](int i, int j) -> const cplx & { return m.Hp[(size_t)i * d + j]; };
    /* the band form must add exactly the dense zgemv's nonzero terms */
    for (int i = 0; i < d; i++)
        for (int j = 0; j < d; j++) {
            const int o = j - i;
            const cplx a = hp(i, j), b = hp(j, i);
            if (a.real() != b.real() || a.imag() != b.imag()) { set_err(e, "H_part is not symmetric"); return -1; }
            if ((o > 2 || o < -2) && (a != 0.0 || m.L[(size_t)i * d + j] != 0.0)) {
                set_err(e, "H_part or L has entries outside the band"); return -1;
            }
            if ((o == 2 || o == -2) && a.real() != 0.0) { set_err(e, "H_part(i, i+2) is not imaginary"); return -1; }
            if ((o == 2 || o == -2) && m.L[(size_t)i * d + j] != 0.0) { set_err(e, "L is not tridiagonal"); return -1; }
            if (o != 0 && m.Hdrv[(size_t)i * d + j] != 0.0) { set_err(e, "H_drv is not diagonal"); return -1; }
            if (m.Hdrv[(size_t)i * d + j].imag() != 0.0 || m.L[(size_t)i * d + j].imag() != 0.0) {
                set_err(e, "H_drv or L is not real"); return -1;
            }
        }
    e->prop = 1;
    e->rk_ns = rk_ns;
    e->drv_type = drv_type;
    e->drv_phase = drv_phase;
    e->rk_step = m.T / double(rk_ns);                     /* rk_proc.cpp:62 */
    e->b_dre.assign(d, 0.0); e->b_dim.assign(d, 0.0); e->b_drv.assign(d, 0.0);
    e->b_o1re.assign(d, 0.0); e->b_o1im.assign(d, 0.0); e->b_o2im.assign(d, 0.0);
    e->b_Ld.assign(d, 0.0); e->b_Lm.assign(d, 0.0); e->b_Lp.assign(d, 0.0);
    for (int i = 0; i < d; i++) {
        e->b_dre[i] = hp(i, i).real();
        e->b_dim[i] = hp(i, i).imag();
        e->b_drv[i] = m.Hdrv[(size_t)i * d + i].real();
        e->b_Ld[i] = m.L[(size_t)i * d + i].real();
        if (i + 1 < d) {
            e->b_o1re[i] = hp(i, i + 1).real();
            e->b_o1im[i] = hp(i, i + 1).imag();
            e->b_Lp[i] = m.L[(size_t)i * d + i + 1].real();
        }
        if (i >= 1) e->b_Lm[i] = m.L[(size_t)i * d + i - 1].real();
        if (i + 2 < d) e->b_o2im[i] = hp(i, i + 2).imag();
    }
    for (int tr = 0; tr < e->M; tr++) e->pos[tr] = 0;
    return 0;
}

extern "C" double qjo_rk_sin(double x) { return rk_sin(x); }

extern "C" double qjo_rk_drive(const qjo_ens *e, double t) { return rk_drive(e, t); }

extern "C" int qjo_rk_rhs(const qjo_ens *e, double t, const double *x, double *k)
{
    if (!e || e->prop != 1) return -1;
    rk_rhs(e, rk_drive(e, t), reinterpret_cast<const cplx *>(x), reinterpret_cast<cplx *>(k));
    return 0;
}

extern "C" double qjo_rk_norm(int d, const double *x)
{
    return rk_norm(reinterpret_cast<const cplx *>(x), d);
}

extern "C" int qjo_rk_int(qjo_ens *e, int tr, double step)
{
    if (!e || e->prop != 1 || tr < 0 || tr >= e->M) return -1;
    RkWork w(e->m.d);
    cplx *phi = &e->psi[(size_t)tr * e->m.d];
    rk_int(e, e->t[tr], step, phi, w.nw.data(), w);
    std::memcpy(phi, w.nw.data(), sizeof(cplx) * e->m.d);
    return 0;
}

extern "C" int qjo_steps(qjo_ens *e, int n, int is_obs)
{
#pragma omp parallel for schedule(dynamic, 1) num_threads(e->threads)
    for (int tr = 0; tr < e->M; tr++) {
        if (e->prop == 1) rk_steps(e, tr, n, is_obs);
        else top_steps(e, tr, n, is_obs);
    }
    return e->failed ? -1 : 0;
}

extern "C" int qjo_periods_one(qjo_ens *e, int tr, int n, int is_obs)
{
    if (tr < 0 || tr >= e->M) return -1;
    if (e->prop == 1) {
        rk_steps(e, tr, n * e->rk_ns, is_obs);
        return e->failed ? -1 : 0;
    }
    for (int p = 0; p < n && !e->failed; p++) one_period(e, tr, is_obs);
    return e->failed ? -1 : 0;
}

extern "C" int qjo_periods(qjo_ens *e, int n, int is_obs)
{
    if (e->prop == 1) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(e->threads)
        for (int tr = 0; tr < e->M; tr++) rk_steps(e, tr, n * e->rk_ns, is_obs);
        return e->failed ? -1 : 0;
    }
    for (int p = 0; p < n && !e->failed; p++) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(e->threads)
        for (int tr = 0; tr < e->M; tr++) one_period(e, tr, is_obs);
    }
    return e->failed ? -1 : 0;
}

extern "C" int qjo_trans(qjo_ens *e, int num_tp_periods)
{
    /* trans_process_single (experiment.cpp:3249-3271), resresh_times (2214-2219) */
#pragma omp parallel for schedule(dynamic, 1) num_threads(e->threads)
    for (int tr = 0; tr < e->M; tr++) {
        double &eta = e->eta[tr];
        eta = 0.0;
        while (eta == 0.0) eta = e->rng[tr].next();
        if (e->prop == 1) {
            /* transient periods are numbered from 0, and so are the observed ones
             * (obser_process starts at dump_periods[0] = 0, experiment.cpp:476-488) */
            e->pos[tr] = 0;
            rk_steps(e, tr, num_tp_periods * e->rk_ns, 0);
            e->pos[tr] = 0;
            continue;
        }
        for (int p = 0; p < num_tp_periods && !e->failed; p++) one_period(e, tr, 0);
    }
    for (int tr = 0; tr < e->M; tr++) e->t[tr] = 0.0;
    return e->failed ? -1 : 0;
}

/* calc_chars_std(_start) (core.cpp:665-725) with get_mean_simple, get_dispersion_simple,
 * get_m2, get_energy (experiment.cpp:1395-1469). */
extern "C" int qjo_chars(qjo_ens *e, int is_start, double *norm, double *mean, double *dispersion,
                         double *m2, double *energy, double *adr_out)
{
    const int d = e->m.d;
    const double E0 = e->p.prm_E;
#pragma omp parallel for schedule(static) num_threads(e->threads)
    for (int tr = 0; tr < e->M; tr++) {
        const cplx *phi = &e->psi[(size_t)tr * d];
        std::vector<double> adr(d);
        const double nrm = norm_square(phi, d);
        for (int k = 0; k < d; k++)
            adr[k] = (phi[k].real() * phi[k].real() + phi[k].imag() * phi[k].imag()) * (1.0 / nrm);
        double mn = 0.0;
        for (int k = 0; k < d; k++) mn += double(k) * adr[k];
        if (is_start) e->mean_start[tr] = mn;
        const double disp = (mn - e->mean_start[tr]) * (mn - e->mean_start[tr]);
        double mm2 = 0.0;
        for (int k = 0; k < d; k++) mm2 += (double(k) - mn) * (double(k) - mn) * adr[k];
        /* energy: <phi|H + E H_drv|phi> / ||phi||^2, dense loop as experiment.cpp:1441-1466 */
        const double sn = std::sqrt(nrm);
        double en = 0.0;
        for (int i = 0; i < d; i++) {
            double sre = 0.0, sim = 0.0;
            for (int j = 0; j < d; j++) {
                const size_t idx = (size_t)i * d + j;
                const double hv = e->m.H[idx].real() + E0 * e->m.Hdrv[idx].real();
                sre += hv * phi[j].real() / sn - 0.0 * phi[j].imag() / sn;
                sim += hv * phi[j].imag() / sn + 0.0 * phi[j].real() / sn;
            }
            en += phi[i].real() / sn * sre + phi[i].imag() / sn * sim;
        }
        if (norm) norm[tr] = nrm;
        if (mean) mean[tr] = mn;
        if (dispersion) dispersion[tr] = disp;
        if (m2) m2[tr] = mm2;
        if (energy) energy[tr] = en;
        if (adr_out) std::memcpy(adr_out + (size_t)tr * d, adr.data(), sizeof(double) * d);
    }
    return 0;
}

extern "C" int qjo_get_state(const qjo_ens *e, double *psi, double *eta, double *t, int *njumps,
                             uint64_t *rng)
{
    const int M = e->M, d = e->m.d;
    if (psi) std::memcpy(psi, e->psi.data(), sizeof(cplx) * (size_t)M * d);
    for (int tr = 0; tr < M; tr++) {
        if (eta) eta[tr] = e->eta[tr];
        if (t) t[tr] = e->t[tr];
        if (njumps) njumps[tr] = e->njumps[tr];
        if (rng) rng[tr] = e->rng[tr].x;
    }
    return 0;
}

extern "C" int qjo_set_state(qjo_ens *e, const double *psi, const double *eta, const double *t,
                             const uint64_t *rng)
{
    const int M = e->M, d = e->m.d;
    if (psi) std::memcpy(e->psi.data(), psi, sizeof(cplx) * (size_t)M * d);
    for (int tr = 0; tr < M; tr++) {
        if (eta) e->eta[tr] = eta[tr];
        if (t) e->t[tr] = t[tr];
        if (rng) e->rng[tr].x = rng[tr];
    }
    return 0;
}

extern "C" int qjo_get_jumps(const qjo_ens *e, int tr, double *times, int *types, double *norms,
                             double *etas, int cap)
{
    if (tr < 0 || tr >= e->M) return -1;
    const int n = (int)e->jt[tr].size();
    for (int i = 0; i < n && i < cap; i++) {
        if (times) times[i] = e->jt[tr][i];
        if (types) types[i] = e->jtype[tr][i];
        if (norms) norms[i] = e->jn[tr][i];
        if (etas) etas[i] = e->je[tr][i];
    }
    return n;
}

extern "C" int64_t qjo_num_matvecs(const qjo_ens *e)
{
    int64_t s = 0;
    for (auto v : e->matvecs) s += v;
    return s;
}

extern "C" int64_t qjo_num_jumps(const qjo_ens *e)
{
    int64_t s = 0;
    for (auto v : e->jumps_done) s += v;
    return s;
}

/* ------------------------------------------------------------------------- */
/* LPN (task_id 7, LpnMultExperimentBehaviour, experiment.cpp:137-286): the   */
/* first M/2 trajectories are the originals, trajectory tr + M/2 is a copy of */
/* tr (state, RNG stream, eta) perturbed to a chosen observable distance, and */
/* the logarithm of the distance growth accumulates into lambda.  Random      */
/* observables: init_random_obs / get_random_obs (newdel.cpp:3652-3722,       */
/* experiment.cpp:2023-2078).                                                  */
/* ------------------------------------------------------------------------- */

/* 10^y (the LPN perturbation size) and ln x (the LPN growth), evaluated the same
 * way here and in the library's device code, where libm's pow and log may differ
 * from the host's in the last place; both stay within 2 ulp of the C library.
 * 10^y = 2^n e^r: t = y ln 10 as a double-double, n the nearest integer to t / ln 2,
 * r = t - n ln 2 with ln 2 split into 32 + 53 significant bits (Cody & Waite), e^r
 * by its Taylor series to degree 14 on |r| <= ln(2)/2 (truncation < 1e-18, Horner).
 * ln x = e ln 2 + 2 atanh s for x = 2^e m, m in [sqrt(1/2), sqrt(2)),
 * s = (m - 1)/(m + 1): the atanh series to s^23 (|s| < 0.172, truncation < 1e-18),
 * arranged as in fdlibm's log so that the leading term m - 1 is exact. */
static double lpn_exp10(double y)
{
    const double th = y * 0x1.26bb1bbb55516p+1;
    const double tl = std::fma(y, 0x1.26bb1bbb55516p+1, -th) + y * -0x1.f48ad494ea3e9p-53;
    const double n = std::nearbyint(th * 0x1.71547652b82fep+0);
    double r = th - n * 0x1.62e42fee00000p-1;
    r = r - n * 0x1.a39ef35793c76p-33;
    r = r + tl;
    double p = 0x1.93974a8c07c9dp-37;
    p = std::fma(p, r, 0x1.6124613a86d09p-33);
    p = std::fma(p, r, 0x1.1eed8eff8d898p-29);
    p = std::fma(p, r, 0x1.ae64567f544e4p-26);
    p = std::fma(p, r, 0x1.27e4fb7789f5cp-22);
    p = std::fma(p, r, 0x1.71de3a556c734p-19);
    p = std::fma(p, r, 0x1.a01a01a01a01ap-16);
    p = std::fma(p, r, 0x1.a01a01a01a01ap-13);
    p = std::fma(p, r, 0x1.6c16c16c16c17p-10);
    p = std::fma(p, r, 0x1.1111111111111p-7);
    p = std::fma(p, r, 0x1.5555555555555p-5);
    p = std::fma(p, r, 0x1.5555555555555p-3);
    p = std::fma(p, r, 0x1.0000000000000p-1);
    p = std::fma(p, r, 1.0);
    p = std::fma(p, r, 1.0);
    return std::ldexp(p, (int)n);
}

static double lpn_log(double x)
{
    int e = 0;
    double m = std::frexp(x, &e);
    if (m < 0x1.6a09e667f3bcdp-1) {
        m = m * 2.0;
        e -= 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    /* 2 atanh s = 2 s + s R, R = sum_k 2 z^k / (2k + 1); with f = 2 s + s f this is
     * f - (f^2/2 - s (f^2/2 + R)), which keeps the leading term f exact */
    double R = 0x1.642c8590b2164p-4;
    R = std::fma(R, z, 0x1.8618618618618p-4);
    R = std::fma(R, z, 0x1.af286bca1af28p-4);
    R = std::fma(R, z, 0x1.e1e1e1e1e1e1ep-4);
    R = std::fma(R, z, 0x1.1111111111111p-3);
    R = std::fma(R, z, 0x1.3b13b13b13b14p-3);
    R = std::fma(R, z, 0x1.745d1745d1746p-3);
    R = std::fma(R, z, 0x1.c71c71c71c71cp-3);
    R = std::fma(R, z, 0x1.2492492492492p-2);
    R = std::fma(R, z, 0x1.999999999999ap-2);
    R = std::fma(R, z, 0x1.5555555555555p-1);
    R = R * z;
    const double hfsq = 0.5 * f * f;
    const double de = double(e);
    return de * 0x1.62e42fee00000p-1 - ((hfsq - (s * (hfsq + R) + de * 0x1.a39ef35793c76p-33)) - f);
}

extern "C" double qjo_lpn_exp10(double y) { return lpn_exp10(y); }
extern "C" double qjo_lpn_log(double x) { return lpn_log(x); }

/* MKL VSL MCG31m1: x <- a x mod (2^31 - 1), a = 1132489760; the [lo, hi) uniform
 * is fma(double(x), (hi - lo) (1 / m), lo) (pinned by tests/golden/mcg31_mkl.json). */
static const uint64_t kMcg31A = 1132489760ULL, kMcg31M = 2147483647ULL;

struct Mcg31 {
    uint64_t x = 1, stride = 1;
    double next(double lo, double hi)
    {
        const double r = std::fma(double(x), (hi - lo) * (1.0 / double(kMcg31M)), lo);
        x = x * stride % kMcg31M;
        return r;
    }
};

static uint64_t mcg31_pow(uint64_t a, uint64_t e)
{
    uint64_t r = 1;
    a %= kMcg31M;
    while (e) {
        if (e & 1ULL) r = r * a % kMcg31M;
        a = a * a % kMcg31M;
        e >>= 1;
    }
    return r;
}

/* vslNewStream(MCG31, seed) + vslLeapfrogStream(k, n) */
static Mcg31 mcg31_stream(uint64_t seed, int k, int n)
{
    Mcg31 s;
    uint64_t x0 = seed % kMcg31M;
    if (x0 == 0) x0 = 1;
    s.x = x0 * mcg31_pow(kMcg31A, (uint64_t)k) % kMcg31M;
    s.stride = mcg31_pow(kMcg31A, (uint64_t)n);
    return s;
}

/* VSL_RNG_METHOD_GAUSSIAN_BOXMULLER: x = sqrt(-2 ln u1) sin(2 pi u2) from two
 * consecutive uniforms (MKL's own evaluation agrees to 2e-8 relative). */
static double box_muller(Mcg31 &s)
{
    const double u1 = s.next(0.0, 1.0);
    const double u2 = s.next(0.0, 1.0);
    return std::sqrt(-2.0 * std::log(u1)) * std::sin(2.0 * kPi * u2);
}

extern "C" void qjo_mcg31_uniform(unsigned seed, int k, int n, int count, double lo, double hi, double *out)
{
    Mcg31 s = mcg31_stream(seed, k, n);
    for (int i = 0; i < count; i++) out[i] = s.next(lo, hi);
}

extern "C" void qjo_mcg31_gaussian(unsigned seed, int k, int n, int count, double *out)
{
    Mcg31 s = mcg31_stream(seed, k, n);
    for (int i = 0; i < count; i++) out[i] = box_muller(s);
}

/* init_random_obs: X = (G_re + i G_im) / 2 with d^2 + d^2 Gaussians of stream
 * leapfrog(seed + obs, mns) of MCG31(3873247); R = X X^dag (ascending k); type > 0
 * divides by Tr R, type 2 by Tr R / d. */
static void random_obs_matrix(int d, int seed, int mns, int type, int obs, std::vector<cplx> &R)
{
    const size_t dd = (size_t)d * d;
    Mcg31 s = mcg31_stream(3873247ULL, seed + obs, mns);
    std::vector<double> gr(dd), gi(dd);
    for (size_t i = 0; i < dd; i++) gr[i] = box_muller(s);
    for (size_t i = 0; i < dd; i++) gi[i] = box_muller(s);
    std::vector<double> xr(dd), xi(dd);
    for (size_t i = 0; i < dd; i++) {
        xr[i] = 0.5 * gr[i];
        xi[i] = 0.5 * gi[i];
    }
    R.assign(dd, 0.0);
    for (int i = 0; i < d; i++)
        for (int j = 0; j < d; j++) {
            double re = 0.0, im = 0.0;
            for (int k = 0; k < d; k++) {
                const double ar = xr[(size_t)i * d + k], ai = xi[(size_t)i * d + k];
                const double br = xr[(size_t)j * d + k], bi = -xi[(size_t)j * d + k];
                re = re + (ar * br - ai * bi);
                im = im + (ar * bi + ai * br);
            }
            R[(size_t)i * d + j] = cplx(re, im);
        }
    if (type > 0) {
        double tr = 0.0;
        for (int i = 0; i < d; i++) tr += R[(size_t)i * d + i].real();
        if (type == 2) tr = tr / double(d);
        for (size_t i = 0; i < dd; i++) R[i] = cplx(R[i].real() / tr, R[i].imag() / tr);
    }
}

extern "C" int qjo_random_obs_matrix(int d, int seed, int mns, int type, int obs, double *out)
{
    std::vector<cplx> R;
    random_obs_matrix(d, seed, mns, type, obs, R);
    std::memcpy(out, R.data(), sizeof(cplx) * R.size());
    return 0;
}

/* get_random_obs: <phi|R|phi> / ||phi||^2 as the reference evaluates it. */
static cplx random_obs_value(const cplx *phi, const cplx *R, int d, std::vector<cplx> &tmp)
{
    const double n = std::sqrt(norm_square(phi, d));
    tmp.resize(2 * (size_t)d);
    cplx *pn = tmp.data(), *mult = tmp.data() + d;
    for (int i = 0; i < d; i++) pn[i] = cplx(phi[i].real() / n, phi[i].imag() / n);
    for (int i = 0; i < d; i++) {
        double re = 0.0, im = 0.0;
        for (int j = 0; j < d; j++) {
            const cplx &a = R[(size_t)i * d + j];
            re = re + (a.real() * pn[j].real() - a.imag() * pn[j].imag());
            im = im + (a.real() * pn[j].imag() + a.imag() * pn[j].real());
        }
        mult[i] = cplx(re, im);
    }
    double rr = 0.0, ri = 0.0;
    for (int i = 0; i < d; i++) {
        const double cr = pn[i].real(), ci = -pn[i].imag();
        rr = rr + (cr * mult[i].real() - ci * mult[i].imag());
        ri = ri + (ci * mult[i].real() + cr * mult[i].imag());
    }
    return cplx(rr, ri);
}

/* calc_random_obs on the STD path (core.cpp:12614-12622, called from calc_chars_std,
 * core.cpp:694, 724): the random observables random_obs_seed + o of every
 * trajectory's current state, M x num complex. */
extern "C" int qjo_random_obs(qjo_ens *e, int num, int seed, int mns, int type, double *out)
{
    const int d = e->m.d, M = e->M;
    for (int o = 0; o < num; o++) {
        std::vector<cplx> R;
        random_obs_matrix(d, seed, mns, type, o, R);
#pragma omp parallel for schedule(static) num_threads(e->threads)
        for (int tr = 0; tr < M; tr++) {
            std::vector<cplx> tmp;
            const cplx v = random_obs_value(&e->psi[(size_t)tr * d], R.data(), d, tmp);
            out[2 * ((size_t)tr * num + o)] = v.real();
            out[2 * ((size_t)tr * num + o) + 1] = v.imag();
        }
    }
    return 0;
}

/* mean and energy of one state as calc_chars_std computes them (the loops of qjo_chars) */
static void mean_energy(const qjo_ens *e, const cplx *phi, double *mean, double *energy)
{
    const int d = e->m.d;
    const double E0 = e->p.prm_E;
    const double nrm = norm_square(phi, d);
    double mn = 0.0;
    for (int k = 0; k < d; k++)
        mn += double(k) * ((phi[k].real() * phi[k].real() + phi[k].imag() * phi[k].imag()) * (1.0 / nrm));
    const double sn = std::sqrt(nrm);
    double en = 0.0;
    for (int i = 0; i < d; i++) {
        double sre = 0.0, sim = 0.0;
        for (int j = 0; j < d; j++) {
            const size_t idx = (size_t)i * d + j;
            const double hv = e->m.H[idx].real() + E0 * e->m.Hdrv[idx].real();
            sre += hv * phi[j].real() / sn - 0.0 * phi[j].imag() / sn;
            sim += hv * phi[j].imag() / sn + 0.0 * phi[j].real() / sn;
        }
        en += phi[i].real() / sn * sre + phi[i].imag() / sn * sim;
    }
    *mean = mn;
    *energy = en;
}

struct Lpn {
    qjo_lpn_desc p{};
    int M = 0, half = 0, period = 0;
    double max_energy = 0.0, T = 0.0;
    std::vector<std::vector<cplx>> R;               /* random observables */
    std::vector<Mcg31> var;                          /* streams_var */
    std::vector<double> mean, energy, mean_lpn, energy_lpn, lambda, lambda_now, delta_s;
    std::vector<int> num_renorms;
    std::vector<std::vector<cplx>> obs, obs_lpn;     /* [tr][obs] */
    std::vector<std::vector<double>> lambdas, deltas_s, deltas_f;
};

static Lpn *lpn_of(qjo_ens *e);

/* calc_chars_std's mean, energy and random observables of trajectory tr */
static void lpn_chars_std(qjo_ens *e, Lpn &L, int tr)
{
    const int d = e->m.d;
    const cplx *phi = &e->psi[(size_t)tr * d];
    mean_energy(e, phi, &L.mean[tr], &L.energy[tr]);
    std::vector<cplx> tmp;
    for (size_t o = 0; o < L.R.size(); o++) L.obs[tr][o] = random_obs_value(phi, L.R[o].data(), d, tmp);
}

/* calc_chars_lpn (core.cpp:761-792) */
static void lpn_chars_lpn(qjo_ens *e, Lpn &L, int tr)
{
    const int d = e->m.d;
    const cplx *phi = &e->psi[(size_t)tr * d];
    mean_energy(e, phi, &L.mean_lpn[tr], &L.energy_lpn[tr]);
    std::vector<cplx> tmp;
    for (size_t o = 0; o < L.R.size(); o++) L.obs_lpn[tr][o] = random_obs_value(phi, L.R[o].data(), d, tmp);
}

/* calc_delta_s (from the lpn values) and calc_delta_f (from the std values),
 * core.cpp:839-915 */
static double lpn_delta(const qjo_ens *e, const Lpn &L, int tr, int base, bool from_lpn)
{
    const int d = e->m.d;
    const std::vector<double> &mn = from_lpn ? L.mean_lpn : L.mean;
    const std::vector<double> &en = from_lpn ? L.energy_lpn : L.energy;
    const std::vector<std::vector<cplx>> &ob = from_lpn ? L.obs_lpn : L.obs;
    if (L.p.lpn_type == -1) return std::fabs(mn[tr] - mn[base]) / double(d);
    if (L.p.lpn_type == -2) return std::fabs(en[tr] - en[base]) / L.max_energy;
    const cplx b = ob[base][L.p.lpn_type], v = ob[tr][L.p.lpn_type];
    const double dr = b.real() - v.real(), di = b.imag() - v.imag();
    const double tmp = dr * dr + di * di;   /* pow(x, 2) is the correctly rounded x * x */
    return std::sqrt(tmp);
}

/* var_trajectory_lpn (experiment.cpp:2278-2393): phi = base + eps var, renormalised
 * to the base norm, eps bisected in log10 between lpn_eps_low and lpn_eps_high
 * until |delta_s - lpn_delta_s| < lpn_eps_error or lpn_eps_deep halvings. */
static void lpn_var(qjo_ens *e, Lpn &L, int tr, int base)
{
    const int d = e->m.d;
    double lo = L.p.lpn_eps_low, hi = L.p.lpn_eps_high, diff = hi - lo;
    int deep = 0;
    double eps = lpn_exp10((lo + hi) * 0.5);
    cplx *phi = &e->psi[(size_t)tr * d];
    const cplx *orig = &e->psi[(size_t)base * d];
    std::vector<double> v(2 * (size_t)d);
    for (int i = 0; i < 2 * d; i++) v[i] = L.var[tr].next(-1.0, 1.0);
    std::vector<cplx> var(d);
    for (int i = 0; i < d; i++) var[i] = cplx(v[i], v[d + i]);
    const double nv2 = norm_square(var.data(), d);
    for (int i = 0; i < d; i++) var[i] = cplx(var[i].real() / std::sqrt(nv2), var[i].imag() / std::sqrt(nv2));
    const double n2 = norm_square(orig, d);
    while (deep < L.p.lpn_eps_deep) {
        for (int i = 0; i < d; i++)
            phi[i] = cplx(orig[i].real() + eps * var[i].real(), orig[i].imag() + eps * var[i].imag());
        const double nm2 = norm_square(phi, d);
        for (int i = 0; i < d; i++) phi[i] = cplx(phi[i].real() / std::sqrt(nm2), phi[i].imag() / std::sqrt(nm2));
        for (int i = 0; i < d; i++) phi[i] = cplx(phi[i].real() * std::sqrt(n2), phi[i].imag() * std::sqrt(n2));
        lpn_chars_lpn(e, L, tr);
        const double ds = lpn_delta(e, L, tr, base, true);
        if (std::fabs(ds - L.p.lpn_delta_s) < L.p.lpn_eps_error) break;
        if (ds > L.p.lpn_delta_s) {
            hi -= 0.5 * diff;
            diff = hi - lo;
            deep++;
            eps = lpn_exp10((lo + hi) * 0.5);
        }
        if (ds < L.p.lpn_delta_s) {
            lo += 0.5 * diff;
            diff = hi - lo;
            deep++;
            eps = lpn_exp10((lo + hi) * 0.5);
        }
        if (!(ds > L.p.lpn_delta_s) && !(ds < L.p.lpn_delta_s)) break;   /* equal or NaN */
    }
}

/* copy_trajectory_lpn / copy_stream_lpn (experiment.cpp:2221-2256) */
static void lpn_copy(qjo_ens *e, int tr, int base, bool state)
{
    const int d = e->m.d;
    e->rng[tr] = e->rng[base];
    e->eta[tr] = e->eta[base];
    if (state) std::memcpy(&e->psi[(size_t)tr * d], &e->psi[(size_t)base * d], sizeof(cplx) * d);
}

/* calc_chars_lpn_start (core.cpp:727-759) */
static void lpn_start(qjo_ens *e, Lpn &L, int tr, int base)
{
    L.lambda[tr] = 0.0;
    L.lambda_now[tr] = 0.0;
    L.mean_lpn[tr] = L.mean[tr];
    L.energy_lpn[tr] = L.energy[tr];
    L.delta_s[tr] = lpn_delta(e, L, tr, base, false);
    L.obs_lpn[tr] = L.obs[tr];
}

extern "C" int qjo_lpn_setup(qjo_ens *e, const qjo_lpn_desc *desc)
{
    if (!e || !desc || e->M < 2 || (e->M % 2)) return -1;
    if (desc->lpn_type < -2 || desc->lpn_type >= desc->num_random_obs) {
        set_err(e, "wrong lpn_type");
        return -1;
    }
    if (!(desc->lpn_eps_error > 0.0)) {
        set_err(e, "lpn_eps_error must be > 0");
        return -1;
    }
    delete e->lpn;
    Lpn *L = new Lpn;
    e->lpn = L;
    L->p = *desc;
    L->M = e->M;
    L->half = e->M / 2;
    L->T = e->m.T;
    const int d = e->m.d, M = e->M;
    for (int i = 0; i < d; i++)                              /* init_obs_lpn, newdel_exp.cpp:540-550 */
        L->max_energy = std::max(L->max_energy, std::fabs(e->m.H[(size_t)i * d + i].real()));
    L->R.resize(desc->num_random_obs);
    for (int o = 0; o < desc->num_random_obs; o++)
        random_obs_matrix(d, desc->random_obs_seed, desc->random_obs_mns, desc->random_obs_type, o, L->R[o]);
    L->var.resize(M);
    for (int tr = 0; tr < M; tr++) L->var[tr] = mcg31_stream(777, tr, M);   /* init_streams_var */
    L->mean.assign(M, 0.0);
    L->energy.assign(M, 0.0);
    L->mean_lpn.assign(M, 0.0);
    L->energy_lpn.assign(M, 0.0);
    L->lambda.assign(M, 0.0);
    L->lambda_now.assign(M, 0.0);
    L->delta_s.assign(M, 0.0);
    L->num_renorms.assign(M, 0);
    L->obs.assign(M, std::vector<cplx>(desc->num_random_obs));
    L->obs_lpn.assign(M, std::vector<cplx>(desc->num_random_obs));
    L->lambdas.assign(M, {});
    L->deltas_s.assign(M, {});
    L->deltas_f.assign(M, {});
    return 0;
}

static Lpn *lpn_of(qjo_ens *e) { return e ? e->lpn : nullptr; }

/* LpnMultExperimentBehaviour::trans_process (experiment.cpp:137-190) */
extern "C" int qjo_lpn_trans(qjo_ens *e, int num_tp_periods)
{
    Lpn *L = lpn_of(e);
    if (!L) return -1;
    const int half = L->half, M = L->M;
#pragma omp parallel for schedule(dynamic, 1) num_threads(e->threads)
    for (int tr = 0; tr < half; tr++) {
        double &eta = e->eta[tr];
        eta = 0.0;
        while (eta == 0.0) eta = e->rng[tr].next();
        if (e->prop == 1) {
            e->pos[tr] = 0;
            rk_steps(e, tr, num_tp_periods * e->rk_ns, 0);
            e->pos[tr] = 0;
        } else {
            for (int p = 0; p < num_tp_periods && !e->failed; p++) one_period(e, tr, 0);
        }
    }
    if (e->failed) return -1;
    for (int tr = 0; tr < half; tr++) {
        lpn_chars_std(e, *L, tr);
        lpn_start(e, *L, tr, tr);
    }
    for (int tr = 0; tr < M; tr++) {
        const int base = tr >= half ? tr - half : tr;
        if (tr >= half) {
            lpn_copy(e, tr, base, true);
            e->pos[tr] = e->pos[base];
            lpn_var(e, *L, tr, base);
        }
        e->t[tr] = 0.0;                                     /* resresh_times */
        lpn_chars_std(e, *L, tr);
        lpn_start(e, *L, tr, base);
    }
    L->period = 0;
    return 0;
}

/* one observation period of LpnMultExperimentBehaviour::obser_process
 * (experiment.cpp:218-252) with lambda_lpn / lambda_lpn_now (experiment.cpp:3065-3134) */
extern "C" int qjo_lpn_period(qjo_ens *e)
{
    Lpn *L = lpn_of(e);
    if (!L) return -1;
    const int half = L->half, M = L->M;
#pragma omp parallel for schedule(dynamic, 1) num_threads(e->threads)
    for (int tr = 0; tr < M; tr++) {
        if (e->prop == 1) rk_steps(e, tr, e->rk_ns, 1);
        else one_period(e, tr, 1);
    }
    if (e->failed) return -1;
    for (int tr = 0; tr < M; tr++) lpn_chars_std(e, *L, tr);
    for (int tr = 0; tr < half; tr++) lpn_chars_lpn(e, *L, tr);
    const double curr_time = double(L->period + 1) * L->T;
    for (int tr = half; tr < M; tr++) {
        const int base = tr - half;
        const double df = lpn_delta(e, *L, tr, base, false);
        const bool renorm = L->p.lambda_per_periods != 0 ||
                            df > std::max(L->p.lpn_delta_f_high, L->delta_s[tr] * 10.0) || df < L->p.lpn_delta_f_low;
        if (renorm) {
            L->lambda[tr] += lpn_log(df / L->delta_s[tr] + 1.0e-16);
            L->num_renorms[tr] += 1;
            if (L->p.save_lambdas > 0) {
                L->lambdas[tr].push_back(df / L->delta_s[tr]);
                L->deltas_s[tr].push_back(L->delta_s[tr]);
                L->deltas_f[tr].push_back(df);
            }
            lpn_copy(e, tr, base, false);
            lpn_var(e, *L, tr, base);
            lpn_chars_lpn(e, *L, tr);
            L->delta_s[tr] = lpn_delta(e, *L, tr, base, true);
            L->lambda_now[tr] = L->lambda[tr] / curr_time;
        } else {
            lpn_chars_lpn(e, *L, tr);
            L->lambda_now[tr] = (L->lambda[tr] + lpn_log(df / L->delta_s[tr] + 1.0e-16)) / curr_time;
        }
    }
    L->period++;
    return 0;
}

extern "C" int qjo_lpn_get(qjo_ens *e, double *lambda_now, double *lambda, int *num_renorms, double *mean_lpn,
                           double *energy_lpn, double *delta_s, double *mean, double *energy)
{
    Lpn *L = lpn_of(e);
    if (!L) return -1;
    for (int tr = 0; tr < L->M; tr++) {
        if (lambda_now) lambda_now[tr] = L->lambda_now[tr];
        if (lambda) lambda[tr] = L->lambda[tr];
        if (num_renorms) num_renorms[tr] = L->num_renorms[tr];
        if (mean_lpn) mean_lpn[tr] = L->mean_lpn[tr];
        if (energy_lpn) energy_lpn[tr] = L->energy_lpn[tr];
        if (delta_s) delta_s[tr] = L->delta_s[tr];
        if (mean) mean[tr] = L->mean[tr];
        if (energy) energy[tr] = L->energy[tr];
    }
    return 0;
}

extern "C" int qjo_lpn_random_obs(qjo_ens *e, double *obs, double *obs_lpn)
{
    Lpn *L = lpn_of(e);
    if (!L) return -1;
    const size_t no = L->R.size();
    for (int tr = 0; tr < L->M; tr++)
        for (size_t o = 0; o < no; o++) {
            if (obs) {
                obs[2 * (tr * no + o)] = L->obs[tr][o].real();
                obs[2 * (tr * no + o) + 1] = L->obs[tr][o].imag();
            }
            if (obs_lpn) {
                obs_lpn[2 * (tr * no + o)] = L->obs_lpn[tr][o].real();
                obs_lpn[2 * (tr * no + o) + 1] = L->obs_lpn[tr][o].imag();
            }
        }
    return 0;
}

extern "C" int qjo_lpn_lambdas(qjo_ens *e, int tr, double *ratio, double *ds, double *df, int cap)
{
    Lpn *L = lpn_of(e);
    if (!L || tr < 0 || tr >= L->M) return -1;
    const int n = (int)L->lambdas[tr].size();
    for (int i = 0; i < n && i < cap; i++) {
        if (ratio) ratio[i] = L->lambdas[tr][i];
        if (ds) ds[i] = L->deltas_s[tr][i];
        if (df) df[i] = L->deltas_f[tr][i];
    }
    return n;
}

qjo_ens::~qjo_ens() { delete lpn; }
