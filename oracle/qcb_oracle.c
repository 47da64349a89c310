/*
 * qcb_oracle.c -- FP64 CPU restatement of ECMC-QCB's Wilson-flow measurement path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: the tests, the
 * smoke check and the CPU-baseline leg of the benchmark call it as the result
 * the GPU path is compared with.
 * Nothing in the shipped library links, loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the ECMC-QCB source tree).  The reference cannot be compiled in this image
 * (Eigen3 and c-lime are absent), so this restatement is pinned by the checks
 * in tests/test_oracle_*.py instead:
 *   exp_su3          -> scipy.linalg.expm golden vectors, closed forms (diagonal,
 *                       SU(2)-embedded), exp(X)exp(-X)=1, det=1
 *   proj_lie_su3     -> idempotence, tracelessness, anti-hermiticity, closed form
 *   staple / force   -> finite-difference identity against the plaquette action,
 *                       gauge covariance, zero force on constant abelian fields
 *   rk3 step         -> O(eps^3) global error at fixed flow time, gauge
 *                       covariance, stationarity of abelian fields
 *   plaquette, clover, Q, E -> closed forms on constant abelian fields,
 *                       gauge / translation invariance, Q -> -Q under reflection
 *   heatbath         -> not a parity target (test-input generator only); checked
 *                       statistically (plaquette at beta=6) and for unitarity.
 *
 * Field layout (same as the ILDG payload and the reference's interior order):
 * site s = ((t*Lz + z)*Ly + y)*Lx + x (x fastest), then mu = 0..3, then the
 * 3x3 matrix row-major, each entry (re, im): 72 doubles per site.  The lattice
 * is a single periodic domain (the reference's halos + exchange_halos_cascade,
 * src/mpi/HalosExchange.cpp:4-72, provide exactly periodic neighbours).
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;
typedef struct {
    cplx m[9];
} mat3; /* row-major 3x3 */

#define M(a, i, j) ((a).m[3 * (i) + (j)])

/* ------------------------------------------------------------------------- */
/* 3x3 complex algebra (replaces Eigen's fixed-size products)                */
/* ------------------------------------------------------------------------- */
static inline mat3 mzero(void) {
    mat3 r;
    for (int k = 0; k < 9; k++) r.m[k] = 0.0;
    return r;
}
static inline mat3 mid(void) {
    mat3 r = mzero();
    r.m[0] = r.m[4] = r.m[8] = 1.0;
    return r;
}
static inline mat3 mmul(mat3 a, mat3 b) {
    mat3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            M(r, i, j) = M(a, i, 0) * M(b, 0, j) + M(a, i, 1) * M(b, 1, j) + M(a, i, 2) * M(b, 2, j);
    return r;
}
static inline mat3 madj(mat3 a) {
    mat3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) M(r, i, j) = conj(M(a, j, i));
    return r;
}
static inline mat3 madd(mat3 a, mat3 b) {
    for (int k = 0; k < 9; k++) a.m[k] += b.m[k];
    return a;
}
static inline mat3 msub(mat3 a, mat3 b) {
    for (int k = 0; k < 9; k++) a.m[k] -= b.m[k];
    return a;
}
static inline mat3 mscale(double c, mat3 a) {
    for (int k = 0; k < 9; k++) a.m[k] *= c;
    return a;
}
static inline cplx mtrace(mat3 a) { return a.m[0] + a.m[4] + a.m[8]; }
static inline cplx mdet(mat3 a) {
    return M(a, 0, 0) * (M(a, 1, 1) * M(a, 2, 2) - M(a, 1, 2) * M(a, 2, 1)) -
           M(a, 0, 1) * (M(a, 1, 0) * M(a, 2, 2) - M(a, 1, 2) * M(a, 2, 0)) +
           M(a, 0, 2) * (M(a, 1, 0) * M(a, 2, 1) - M(a, 1, 1) * M(a, 2, 0));
}
static inline double mnorm2(mat3 a) {
    double s = 0.0;
    for (int k = 0; k < 9; k++) s += creal(a.m[k]) * creal(a.m[k]) + cimag(a.m[k]) * cimag(a.m[k]);
    return s;
}

static inline mat3 load(const double* p) {
    mat3 r;
    for (int k = 0; k < 9; k++) r.m[k] = CMPLX(p[2 * k], p[2 * k + 1]);
    return r;
}
static inline void store(double* p, mat3 a) {
    for (int k = 0; k < 9; k++) {
        p[2 * k] = creal(a.m[k]);
        p[2 * k + 1] = cimag(a.m[k]);
    }
}

/* ------------------------------------------------------------------------- */
/* Geometry: periodic lexicographic index, x fastest.                        */
/* Restates GeometryCB::index (src/geometry/GeometryCB.h:25-27) and the      */
/* neighbour table (src/geometry/GeometryCB.cpp:10-31) on a periodic domain. */
/* ------------------------------------------------------------------------- */
typedef struct {
    int L[4];
    long V;
    long* nb; /* nb[s*8 + mu*2 + d], d = 0 up, 1 down (GeometryCB.h:30-32) */
} geom;

static void geom_init(geom* g, const int L[4]) {
    for (int d = 0; d < 4; d++) g->L[d] = L[d];
    g->V = (long)L[0] * L[1] * L[2] * L[3];
    g->nb = (long*)malloc(sizeof(long) * 8 * g->V);
    for (int t = 0; t < L[3]; t++)
        for (int z = 0; z < L[2]; z++)
            for (int y = 0; y < L[1]; y++)
                for (int x = 0; x < L[0]; x++) {
                    int c[4] = {x, y, z, t};
                    long s = (((long)t * L[2] + z) * L[1] + y) * L[0] + x;
                    for (int mu = 0; mu < 4; mu++) {
                        for (int d = 0; d < 2; d++) {
                            int cc[4] = {c[0], c[1], c[2], c[3]};
                            cc[mu] = (cc[mu] + (d == 0 ? 1 : L[mu] - 1)) % L[mu];
                            g->nb[s * 8 + mu * 2 + d] =
                                (((long)cc[3] * L[2] + cc[2]) * L[1] + cc[1]) * L[0] + cc[0];
                        }
                    }
                }
}
static void geom_free(geom* g) { free(g->nb); }
#define UP(g, s, mu) ((g)->nb[(s)*8 + (mu)*2 + 0])
#define DN(g, s, mu) ((g)->nb[(s)*8 + (mu)*2 + 1])

static inline mat3 link(const double* f, long s, int mu) { return load(f + (s * 4 + mu) * 18); }
static inline void set_link(double* f, long s, int mu, mat3 a) { store(f + (s * 4 + mu) * 18, a); }

/* ------------------------------------------------------------------------- */
/* su(3) projection and exponential                                          */
/* ------------------------------------------------------------------------- */

/* proj_lie_su3: src/su3/utils.cpp:130-134.  B = (A - A^dag)/2; A = B - tr(B)/3. */
static mat3 proj_lie_su3(mat3 a) {
    mat3 b = mscale(0.5, msub(a, madj(a)));
    cplx tr3 = mtrace(b) / 3.0;
    b.m[0] -= tr3;
    b.m[4] -= tr3;
    b.m[8] -= tr3;
    return b;
}

static int g_exp_warnings = 0;

/* exp_su3_luscher: src/su3/utils.cpp:145-193.
 * X = coeff*Z; t = -1/2 Re tr X^2 (:151); d = Re(i det X) (:152); warn if t > 1
 * (:153-155); c_n = 1/n! by repeated division, N = 18 (:158-166); Horner
 * q0' = c_n - i d q2, q1' = q0 - t q2, q2' = q1 for n = 17..0 (:172-186);
 * result q0 1 + q1 X + q2 X^2 (:189-192). */
static mat3 exp_su3(mat3 z, double coeff) {
    mat3 x = mscale(coeff, z);
    mat3 x2 = mmul(x, x);
    double t = -0.5 * creal(mtrace(x2));
    double d = creal(I * mdet(x));
    if (t > 1.0) g_exp_warnings++;
    enum { N = 18 };
    double c[N + 1];
    c[0] = 1.0;
    for (int n = 1; n <= N; n++) c[n] = c[n - 1] / (double)n;
    cplx q0 = c[N], q1 = 0.0, q2 = 0.0;
    for (int n = N - 1; n >= 0; n--) {
        cplx q0n = c[n] - I * d * q2;
        cplx q1n = q0 - t * q2;
        cplx q2n = q1;
        q0 = q0n;
        q1 = q1n;
        q2 = q2n;
    }
    mat3 r = mzero();
    for (int k = 0; k < 9; k++) r.m[k] = q1 * x.m[k] + q2 * x2.m[k];
    r.m[0] += q0;
    r.m[4] += q0;
    r.m[8] += q0;
    return r;
}

/* ------------------------------------------------------------------------- */
/* Staples and force                                                         */
/* ------------------------------------------------------------------------- */

/* GaugeField::compute_staple: src/gauge/GaugeField.cpp:74-97.
 * forward  U_nu(x+mu) (U_nu(x) U_mu(x+nu))^dag
 * backward (U_mu(x-nu) U_nu(x+mu-nu))^dag U_nu(x-nu)                     */
static mat3 staple(const double* f, const geom* g, long x, int mu) {
    mat3 st = mzero();
    long xmu = UP(g, x, mu);
    for (int nu = 0; nu < 4; nu++) {
        if (nu == mu) continue;
        long xnu = UP(g, x, nu);
        mat3 U0 = link(f, xmu, nu), U1 = link(f, xnu, mu), U2 = link(f, x, nu);
        st = madd(st, mmul(U0, madj(mmul(U2, U1))));
        long xmunu = DN(g, xmu, nu);
        long xmnu = DN(g, x, nu);
        mat3 V0 = link(f, xmunu, nu), V1 = link(f, xmnu, mu), V2 = link(f, xmnu, nu);
        st = madd(st, mmul(madj(mmul(V1, V0)), V2));
    }
    return st;
}

/* GradientFlow::compute_force: src/flow/gradient_flow.cpp:17-37.
 * F(x,mu) = P_TA(-U_mu(x) Sigma_mu(x)) for every site and direction. */
static void compute_force(const double* f, const geom* g, double* force) {
#pragma omp parallel for schedule(static)
    for (long s = 0; s < g->V; s++)
        for (int mu = 0; mu < 4; mu++) {
            mat3 st = staple(f, g, s, mu);
            mat3 tmp = mscale(-1.0, mmul(link(f, s, mu), st));
            set_link(force, s, mu, proj_lie_su3(tmp));
        }
}

/* compute_w1: src/flow/gradient_flow.cpp:40-57.  U <- exp(eps/4 F0) U. */
static void compute_w1(double* f, const geom* g, const double* F0, double eps) {
#pragma omp parallel for schedule(static)
    for (long s = 0; s < g->V; s++)
        for (int mu = 0; mu < 4; mu++)
            set_link(f, s, mu, mmul(exp_su3(link(F0, s, mu), 0.25 * eps), link(f, s, mu)));
}

/* compute_w2: src/flow/gradient_flow.cpp:60-79.
 * Z = (8/9) eps F1 - (17/36) eps F0; U <- exp(Z) U. */
static void compute_w2(double* f, const geom* g, const double* F0, const double* F1, double eps) {
#pragma omp parallel for schedule(static)
    for (long s = 0; s < g->V; s++)
        for (int mu = 0; mu < 4; mu++) {
            mat3 Z = msub(mscale((8.0 / 9.0) * eps, link(F1, s, mu)),
                          mscale((17.0 / 36.0) * eps, link(F0, s, mu)));
            set_link(f, s, mu, mmul(exp_su3(Z, 1.0), link(f, s, mu)));
        }
}

/* replace_force_0: src/flow/gradient_flow.cpp:82-99.  F0 <- 8/9 F1 - 17/36 F0. */
static void replace_force_0(const geom* g, double* F0, const double* F1) {
#pragma omp parallel for schedule(static)
    for (long s = 0; s < g->V; s++)
        for (int mu = 0; mu < 4; mu++)
            set_link(F0, s, mu,
                     msub(mscale(8.0 / 9.0, link(F1, s, mu)), mscale(17.0 / 36.0, link(F0, s, mu))));
}

/* compute_w3: src/flow/gradient_flow.cpp:102-121.
 * Z = 3/4 eps F1 - eps F0; U <- exp(Z) U. */
static void compute_w3(double* f, const geom* g, const double* F0, const double* F1, double eps) {
#pragma omp parallel for schedule(static)
    for (long s = 0; s < g->V; s++)
        for (int mu = 0; mu < 4; mu++) {
            mat3 Z = msub(mscale(0.75 * eps, link(F1, s, mu)), mscale(eps, link(F0, s, mu)));
            set_link(f, s, mu, mmul(exp_su3(Z, 1.0), link(f, s, mu)));
        }
}

/* rk3_step: src/flow/gradient_flow.cpp:124-135 (halo exchanges are implicit:
 * the domain is periodic).  F0/F1 are scratch fields of V*72 doubles. */
static void rk3_step(double* f, const geom* g, double eps, double* F0, double* F1) {
    compute_force(f, g, F0);
    compute_w1(f, g, F0, eps);
    compute_force(f, g, F1);
    compute_w2(f, g, F0, F1, eps);
    replace_force_0(g, F0, F1);
    compute_force(f, g, F1);
    compute_w3(f, g, F0, F1, eps);
}

/* ------------------------------------------------------------------------- */
/* Observables                                                               */
/* ------------------------------------------------------------------------- */

/* Deterministic reduction: partial sums per (t,z) row block, then an ordered
 * sum of the partials, so results do not depend on the thread count. */

/* mean_plaquette_local/global: src/observables/observables_mpi.cpp:11-48.
 * P = sum_x sum_{mu<nu} Re tr(U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag)/3 / (6V) */
static double plaquette(const double* f, const geom* g) {
    long nrow = (long)g->L[3] * g->L[2];
    long rowlen = (long)g->L[1] * g->L[0];
    double* part = (double*)calloc(nrow, sizeof(double));
#pragma omp parallel for schedule(static)
    for (long r = 0; r < nrow; r++) {
        double sum = 0.0;
        for (long k = 0; k < rowlen; k++) {
            long s = r * rowlen + k;
            for (int mu = 0; mu < 4; mu++)
                for (int nu = mu + 1; nu < 4; nu++) {
                    mat3 U1 = link(f, s, mu);
                    mat3 U2 = link(f, UP(g, s, mu), nu);
                    mat3 U3 = madj(link(f, UP(g, s, nu), mu));
                    mat3 U4 = madj(link(f, s, nu));
                    sum += creal(mtrace(mmul(mmul(mmul(U1, U2), U3), U4))) / 3.0;
                }
        }
        part[r] = sum;
    }
    double tot = 0.0;
    for (long r = 0; r < nrow; r++) tot += part[r];
    free(part);
    return tot / (6.0 * (double)g->V);
}

/* clover_site: src/observables/observables_mpi.cpp:66-90.
 * Four leaves, all anticlockwise in the (mu,nu) plane; F = (C - C^dag)/8. */
static mat3 clover(const double* f, const geom* g, long x, int mu, int nu) {
    long xpmu = UP(g, x, mu), xpnu = UP(g, x, nu);
    long xmmu = DN(g, x, mu), xmnu = DN(g, x, nu);
    long xmmupnu = UP(g, xmmu, nu), xmmumnu = DN(g, xmmu, nu), xpmumnu = DN(g, xpmu, nu);
    mat3 C = mzero();
    C = madd(C, mmul(mmul(mmul(link(f, x, mu), link(f, xpmu, nu)), madj(link(f, xpnu, mu))),
                     madj(link(f, x, nu))));
    C = madd(C, mmul(mmul(mmul(link(f, x, nu), madj(link(f, xmmupnu, mu))), madj(link(f, xmmu, nu))),
                     link(f, xmmu, mu)));
    C = madd(C, mmul(mmul(mmul(madj(link(f, xmmu, mu)), madj(link(f, xmmumnu, nu))),
                          link(f, xmmumnu, mu)),
                     link(f, xmnu, nu)));
    C = madd(C, mmul(mmul(mmul(madj(link(f, xmnu, nu)), link(f, xmnu, mu)), link(f, xpmumnu, nu)),
                     madj(link(f, x, mu))));
    return mscale(0.125, msub(C, madj(C)));
}

/* local_q_e_clover / topo_q_e_clover / topo_q_e_clover_global:
 * src/observables/observables_mpi.cpp:93-156.
 * q = [Re tr F01F23 - Re tr F02F13 + Re tr F03F12] / (4 pi^2);
 * e = 1/2 sum_{mu<nu} |F_munu|^2;  returns {sum q, sum e / V}. */
/* local_q_e_clover (observables_mpi.cpp:93-112) at site s: q(s) with its
 * 1/(4 pi^2) and e(s). */
static void local_q_e(const double* f, const geom* g, long s, double* q_out, double* e_out) {
    mat3 F01 = clover(f, g, s, 0, 1), F02 = clover(f, g, s, 0, 2);
    mat3 F03 = clover(f, g, s, 0, 3), F12 = clover(f, g, s, 1, 2);
    mat3 F13 = clover(f, g, s, 1, 3), F23 = clover(f, g, s, 2, 3);
    double q = creal(mtrace(mmul(F01, F23))) - creal(mtrace(mmul(F02, F13))) +
               creal(mtrace(mmul(F03, F12)));
    double e = 0.5 * (mnorm2(F01) + mnorm2(F02) + mnorm2(F03) + mnorm2(F12) + mnorm2(F13) +
                      mnorm2(F23));
    *q_out = q * (1.0 / (4.0 * M_PI * M_PI));
    *e_out = e;
}

static void q_e_clover(const double* f, const geom* g, double* Q, double* E) {
    long nrow = (long)g->L[3] * g->L[2];
    long rowlen = (long)g->L[1] * g->L[0];
    double* pq = (double*)calloc(nrow, sizeof(double));
    double* pe = (double*)calloc(nrow, sizeof(double));
#pragma omp parallel for schedule(static)
    for (long r = 0; r < nrow; r++) {
        double sq = 0.0, se = 0.0;
        for (long k = 0; k < rowlen; k++) {
            double q, e;
            local_q_e(f, g, r * rowlen + k, &q, &e);
            sq += q;
            se += e;
        }
        pq[r] = sq;
        pe[r] = se;
    }
    double tq = 0.0, te = 0.0;
    for (long r = 0; r < nrow; r++) {
        tq += pq[r];
        te += pe[r];
    }
    free(pq);
    free(pe);
    *Q = tq;
    *E = te / (double)g->V;
}

/* ------------------------------------------------------------------------- */
/* Test-input generation (not a parity target): Haar-random SU(3) hot start   */
/* and a Cabibbo-Marinari / Kennedy-Pendleton heatbath.                       */
/* Restates src/su3/utils.cpp:13-30 (random_su3), :33-63 (SU(2) helpers),    */
/* src/heatbath/heatbath.cpp:8-44 and src/heatbath/heatbath_mpi.cpp:10-33,    */
/* with a portable xoshiro256** stream and a race-free sweep order (one       */
/* direction and one site parity at a time), so configurations are            */
/* reproducible on any host.  They are not bit-identical to the reference's.  */
/* ------------------------------------------------------------------------- */
typedef struct {
    uint64_t s[4];
} rng_t;
static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static inline uint64_t splitmix(uint64_t* x) {
    uint64_t z = (*x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
static void rng_seed(rng_t* r, uint64_t seed) {
    for (int k = 0; k < 4; k++) r->s[k] = splitmix(&seed);
}
static inline uint64_t rng_next(rng_t* r) {
    uint64_t* s = r->s;
    uint64_t res = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return res;
}
/* uniform in (0,1]: never 0, so log() is finite */
static inline double rng_unif(rng_t* r) { return ((rng_next(r) >> 11) + 1) * 0x1.0p-53; }
static inline double rng_gauss(rng_t* r) {
    double u1 = rng_unif(r), u2 = rng_unif(r);
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

/* Haar-random SU(3): complex Gaussian matrix -> column Gram-Schmidt (a QR with
 * positive R diagonal, hence Haar on U(3)) -> divide by det^{1/3}. */
static mat3 random_su3(rng_t* r) {
    mat3 z;
    for (int k = 0; k < 9; k++) {
        double re = rng_gauss(r), im = rng_gauss(r);
        z.m[k] = CMPLX(re, im);
    }
    for (int j = 0; j < 3; j++) {
        for (int p = 0; p < j; p++) {
            cplx dot = 0.0;
            for (int i = 0; i < 3; i++) dot += conj(M(z, i, p)) * M(z, i, j);
            for (int i = 0; i < 3; i++) M(z, i, j) -= dot * M(z, i, p);
        }
        double n = 0.0;
        for (int i = 0; i < 3; i++) n += creal(M(z, i, j) * conj(M(z, i, j)));
        n = sqrt(n);
        for (int i = 0; i < 3; i++) M(z, i, j) /= n;
    }
    cplx det = mdet(z);
    cplx root = cpow(det, 1.0 / 3.0);
    for (int k = 0; k < 9; k++) z.m[k] /= root;
    return z;
}

/* heatbath::generate_su2_kp: src/heatbath/heatbath.cpp:8-36 */
static void su2_kp(double beta, double k, rng_t* r, double a[4]) {
    double alpha = beta * k * 2.0 / 3.0;
    double a0;
    for (;;) {
        double delta = rng_unif(r), r1 = rng_unif(r), r2 = rng_unif(r), r3 = rng_unif(r);
        double cs = cos(2 * M_PI * r2);
        double lsq = -0.5 / alpha * (log(r1) + cs * cs * log(r3));
        if (delta * delta <= 1.0 - lsq) {
            a0 = 1.0 - 2.0 * lsq;
            break;
        }
    }
    double rs = sqrt(fmax(0.0, 1.0 - a0 * a0));
    double phi = 2.0 * M_PI * rng_unif(r);
    double ct = 2.0 * rng_unif(r) - 1.0;
    double st = sqrt(fmax(0.0, 1.0 - ct * ct));
    a[0] = a0;
    a[1] = rs * st * cos(phi);
    a[2] = rs * st * sin(phi);
    a[3] = rs * ct;
}

/* One Cabibbo-Marinari hit on U_mu(x): src/heatbath/heatbath_mpi.cpp:10-33,
 * su2_to_quaternion / su2_quaternion_to_su3 / mult / adj: src/su3/utils.cpp:33-63,
 * su2_step: src/heatbath/heatbath.cpp:39-44. */
static void hb_hit(double* f, const geom* g, long s, int mu, double beta, rng_t* r) {
    mat3 A = staple(f, g, s, mu);
    static const int sub[3][2] = {{0, 1}, {1, 2}, {0, 2}};
    for (int k = 0; k < 3; k++) {
        int i = sub[k][0], j = sub[k][1];
        mat3 W = mmul(link(f, s, mu), A);
        cplx w00 = M(W, i, i), w01 = M(W, i, j), w10 = M(W, j, i), w11 = M(W, j, j);
        double w[4] = {creal(w00 + w11) / 2.0, cimag(w01 + w10) / 2.0, creal(w01 - w10) / 2.0,
                       cimag(w00 - w11) / 2.0};
        double kn = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2] + w[3] * w[3]);
        double v[4] = {w[0] / kn, -w[1] / kn, -w[2] / kn, -w[3] / kn}; /* adj(w/k) */
        double x[4];
        su2_kp(beta, kn, r, x);
        double q[4] = {x[0] * v[0] - x[1] * v[1] - x[2] * v[2] - x[3] * v[3],
                       x[0] * v[1] + x[1] * v[0] + x[2] * v[3] - x[3] * v[2],
                       x[0] * v[2] - x[1] * v[3] + x[2] * v[0] + x[3] * v[1],
                       x[0] * v[3] + x[1] * v[2] - x[2] * v[1] + x[3] * v[0]};
        mat3 R = mzero();
        int kk = 3 - i - j;
        M(R, kk, kk) = 1.0;
        M(R, i, i) = CMPLX(q[0], q[3]);
        M(R, j, j) = CMPLX(q[0], -q[3]);
        M(R, i, j) = CMPLX(q[2], q[1]);
        M(R, j, i) = CMPLX(-q[2], q[1]);
        set_link(f, s, mu, mmul(R, link(f, s, mu)));
    }
}

/* Reunitarise by column Gram-Schmidt (GaugeField::projection_su3,
 * src/gauge/GaugeField.cpp:36-58): c0 normalised, c1 orthogonalised and
 * normalised, c2 = conj(c0 x c1). */
static mat3 proj_su3(mat3 a) {
    cplx c0[3], c1[3], c2[3];
    double n = 0.0;
    for (int i = 0; i < 3; i++) {
        c0[i] = M(a, i, 0);
        n += creal(c0[i] * conj(c0[i]));
    }
    n = sqrt(n);
    for (int i = 0; i < 3; i++) c0[i] /= n;
    cplx dot = 0.0;
    for (int i = 0; i < 3; i++) {
        c1[i] = M(a, i, 1);
        dot += conj(c0[i]) * c1[i];
    }
    n = 0.0;
    for (int i = 0; i < 3; i++) {
        c1[i] -= c0[i] * dot;
        n += creal(c1[i] * conj(c1[i]));
    }
    n = sqrt(n);
    for (int i = 0; i < 3; i++) c1[i] /= n;
    c2[0] = conj(c0[1] * c1[2] - c0[2] * c1[1]);
    c2[1] = conj(c0[2] * c1[0] - c0[0] * c1[2]);
    c2[2] = conj(c0[0] * c1[1] - c0[1] * c1[0]);
    mat3 r;
    for (int i = 0; i < 3; i++) {
        M(r, i, 0) = c0[i];
        M(r, i, 1) = c1[i];
        M(r, i, 2) = c2[i];
    }
    return r;
}

/* ========================================================================= */
/* Exported C API (ctypes).                                                  */
/* ========================================================================= */

int orc_exp_warnings(void) { return g_exp_warnings; }
void orc_reset_exp_warnings(void) { g_exp_warnings = 0; }

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void orc_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* out = exp_su3_luscher(Z, coeff) for n matrices (18 doubles each) */
void orc_exp_su3(const double* Z, int n, double coeff, double* out) {
    for (int k = 0; k < n; k++) store(out + 18 * k, exp_su3(load(Z + 18 * k), coeff));
}
void orc_proj_lie_su3(const double* A, int n, double* out) {
    for (int k = 0; k < n; k++) store(out + 18 * k, proj_lie_su3(load(A + 18 * k)));
}
void orc_proj_su3(const double* A, int n, double* out) {
    for (int k = 0; k < n; k++) store(out + 18 * k, proj_su3(load(A + 18 * k)));
}

void orc_staple(const double* f, const int L[4], long site, int mu, double* out) {
    geom g;
    geom_init(&g, L);
    store(out, staple(f, &g, site, mu));
    geom_free(&g);
}
void orc_compute_force(const double* f, const int L[4], double* force) {
    geom g;
    geom_init(&g, L);
    compute_force(f, &g, force);
    geom_free(&g);
}

/* n RK3 steps in place */
void orc_rk3(double* f, const int L[4], double eps, int n) {
    geom g;
    geom_init(&g, L);
    double* F0 = (double*)malloc(sizeof(double) * 72 * g.V);
    double* F1 = (double*)malloc(sizeof(double) * 72 * g.V);
    for (int k = 0; k < n; k++) rk3_step(f, &g, eps, F0, F1);
    free(F0);
    free(F1);
    geom_free(&g);
}

double orc_plaquette(const double* f, const int L[4]) {
    geom g;
    geom_init(&g, L);
    double p = plaquette(f, &g);
    geom_free(&g);
    return p;
}
void orc_clover(const double* f, const int L[4], long site, int mu, int nu, double* out) {
    geom g;
    geom_init(&g, L);
    store(out, clover(f, &g, site, mu, nu));
    geom_free(&g);
}
void orc_local_q_e(const double* f, const int L[4], long site, double* out) {
    geom g;
    geom_init(&g, L);
    local_q_e(f, &g, site, &out[0], &out[1]);
    geom_free(&g);
}
void orc_q_e_clover(const double* f, const int L[4], double* Q, double* E) {
    geom g;
    geom_init(&g, L);
    q_e_clover(f, &g, Q, E);
    geom_free(&g);
}

/* topo_charge_flowed: src/observables/observables_mpi.cpp:159-186.
 * The input field is copied (gf.copy, :164) and left unmodified.  For
 * n = 0..n_steps_gf-1: measure (t, Q, E) at t; if not last, n_rk_steps RK3
 * steps and t += eps each.  tQE gets 3*n_steps_gf doubles.  The plaquette of
 * each measured field is written to plaq[n] when plaq is non-NULL (the
 * reference does not measure it along the flow).  The flowed field is written
 * to out_field when non-NULL. */
void orc_flow_measure(const double* field, const int L[4], double eps, int n_steps_gf,
                      int n_rk_steps, double* tQE, double* plaq, double* out_field) {
    geom g;
    geom_init(&g, L);
    double* f = (double*)malloc(sizeof(double) * 72 * g.V);
    memcpy(f, field, sizeof(double) * 72 * g.V);
    double* F0 = (double*)malloc(sizeof(double) * 72 * g.V);
    double* F1 = (double*)malloc(sizeof(double) * 72 * g.V);
    double t = 0.0;
    for (int n = 0; n < n_steps_gf; n++) {
        double Q, E;
        q_e_clover(f, &g, &Q, &E);
        tQE[3 * n] = t;
        tQE[3 * n + 1] = Q;
        tQE[3 * n + 2] = E;
        if (plaq) plaq[n] = plaquette(f, &g);
        if (n != n_steps_gf - 1) {
            for (int k = 0; k < n_rk_steps; k++) {
                rk3_step(f, &g, eps, F0, F1);
                t += eps;
            }
        }
    }
    if (out_field) memcpy(out_field, f, sizeof(double) * 72 * g.V);
    free(F0);
    free(F1);
    free(f);
    geom_free(&g);
}

/* Timed RK3 loop (CPU baseline): n_steps RK3 steps, each followed by one
 * measurement of (Q, E, P) when measure != 0.  Returns wall seconds of the
 * loop (omp_get_wtime, or clock_gettime without OpenMP). */
#include <time.h>
double orc_time_flow(double* f, const int L[4], double eps, int n_steps, int measure) {
    geom g;
    geom_init(&g, L);
    double* F0 = (double*)malloc(sizeof(double) * 72 * g.V);
    double* F1 = (double*)malloc(sizeof(double) * 72 * g.V);
    struct timespec a, b;
    clock_gettime(CLOCK_MONOTONIC, &a);
    for (int k = 0; k < n_steps; k++) {
        rk3_step(f, &g, eps, F0, F1);
        if (measure) {
            double Q, E;
            q_e_clover(f, &g, &Q, &E);
            volatile double p = plaquette(f, &g);
            (void)p;
        }
    }
    clock_gettime(CLOCK_MONOTONIC, &b);
    free(F0);
    free(F1);
    geom_free(&g);
    return (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec);
}

/* Test-input generation. */
void orc_hot_start(double* f, const int L[4], uint64_t seed) {
    rng_t r;
    rng_seed(&r, seed);
    long V = (long)L[0] * L[1] * L[2] * L[3];
    for (long s = 0; s < V; s++)
        for (int mu = 0; mu < 4; mu++) set_link(f, s, mu, random_su3(&r));
}
void orc_cold_start(double* f, const int L[4]) {
    long V = (long)L[0] * L[1] * L[2] * L[3];
    for (long s = 0; s < V; s++)
        for (int mu = 0; mu < 4; mu++) set_link(f, s, mu, mid());
}
/* n_sweeps heatbath sweeps.  Each sweep visits, for mu = 0..3 and parity
 * even then odd, every U_mu(x) with that site parity (such links share no
 * staple, so the order inside a pass is irrelevant), then reunitarises. */
void orc_heatbath(double* f, const int L[4], double beta, int n_sweeps, uint64_t seed) {
    geom g;
    geom_init(&g, L);
    rng_t r;
    rng_seed(&r, seed);
    for (int sw = 0; sw < n_sweeps; sw++) {
        for (int mu = 0; mu < 4; mu++)
            for (int par = 0; par < 2; par++)
                for (long s = 0; s < g.V; s++) {
                    long rem = s;
                    int x = rem % L[0];
                    rem /= L[0];
                    int y = rem % L[1];
                    rem /= L[1];
                    int z = rem % L[2];
                    int t = (int)(rem / L[2]);
                    if (((x + y + z + t) & 1) != par) continue;
                    hb_hit(f, &g, s, mu, beta, &r);
                }
        for (long s = 0; s < g.V; s++)
            for (int mu = 0; mu < 4; mu++) set_link(f, s, mu, proj_su3(link(f, s, mu)));
    }
    geom_free(&g);
}
/* random gauge transformation g(x) (Haar), U_mu(x) -> g(x) U_mu(x) g(x+mu)^dag,
 * as GaugeField::gauge_transform (src/gauge/GaugeField.cpp:100-114).  The
 * transformation matrices are returned in gmat (V*18 doubles) if non-NULL. */
void orc_gauge_transform(double* f, const int L[4], uint64_t seed, double* gmat) {
    geom g;
    geom_init(&g, L);
    rng_t r;
    rng_seed(&r, seed);
    mat3* G = (mat3*)malloc(sizeof(mat3) * g.V);
    for (long s = 0; s < g.V; s++) G[s] = random_su3(&r);
    for (long s = 0; s < g.V; s++)
        for (int mu = 0; mu < 4; mu++)
            set_link(f, s, mu, mmul(mmul(G[s], link(f, s, mu)), madj(G[UP(&g, s, mu)])));
    if (gmat)
        for (long s = 0; s < g.V; s++) store(gmat + 18 * s, G[s]);
    free(G);
    geom_free(&g);
}
