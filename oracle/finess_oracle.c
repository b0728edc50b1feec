/*
 * finess_oracle.c -- CPU restatement of FINESS's finite-difference WENO
 * operator and time steppers (the checker for the HIP path; TEST
 * INFRASTRUCTURE ONLY -- nothing in the product links or calls this file).
 *
 * It follows the reference's loop order, per-interface callback granularity
 * and floating-point operation order (pow() where the reference calls pow),
 * with OpenMP over the slow index as the reference does.  Each function cites
 * the reference lines it restates.  Differences from the reference:
 *   - no per-interface heap temporaries (the reference allocates ~20 dTensor
 *     objects per interface, lib/tensors.cpp:188-196): stack arrays instead;
 *   - error paths return status codes instead of exit(1);
 *   - ghost entries of Lstar are zero instead of uninitialized, and corner
 *     ghosts are always defined (the reference's transmissive BC leaves them
 *     unset, apps/2d/euler/riemann_test/SetBndValues.cpp:114);
 *   - Euler Jacobian/Hessian are evaluated exactly with hyper-dual numbers
 *     instead of the closed forms of apps/2d/euler/lib/{DFluxFunc,D2FluxFunc}.cpp;
 *     both are the derivatives of FluxFunc.  They coincide wherever rho*w == 0;
 *     for rho*w != 0 the reference's DFluxFunc has terms pow(q4/q1,3) that
 *     differ from the derivative (DFluxFunc.cpp:43,62,68).
 */
#include "finess_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXEQ 5

/* ---------------------------------------------------------------------- */
/* layout: dTensorBC3(s1,s2,s3,mbc) with ghosts on the first two indices   */
/* (lib/tensors.cpp:650-667), 1-based.                                      */
/* ---------------------------------------------------------------------- */
static inline size_t ix3(int i, int j, int m, int s2, int s3, int mbc)
{
    return ((size_t)(i + mbc - 1) * (size_t)(s2 + 2 * mbc) + (size_t)(j + mbc - 1)) * (size_t)s3
           + (size_t)(m - 1);
}
static inline size_t ix2(int i, int m, int s2, int mbc)
{
    return (size_t)(i + mbc - 1) * (size_t)s2 + (size_t)(m - 1);
}
static inline double Max(double a, double b) { return a > b ? a : b; } /* lib/dog_math.h:18-22 */
static inline double Min(double a, double b) { return a < b ? a : b; }

static size_t state_numel(const finess_params* p, int fields)
{
    if (p->ndims == 1) return (size_t)(p->mx + 2 * p->mbc) * (size_t)fields;
    return (size_t)(p->mx + 2 * p->mbc) * (size_t)(p->my + 2 * p->mbc) * (size_t)fields;
}
static int naux(const finess_params* p) { return p->maux > 0 ? p->maux : 1; } /* StateVars.h:9-53 */

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void orc_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ====================================================================== */
/* WENO reconstructions                                                     */
/* ====================================================================== */

/* WenoReconstruct_JS5, lib/WenoReconstruct.cpp:90-141 */
static double weno_js5(double eps, double pw, const double* g)
{
    const double uim2 = g[0], uim1 = g[1], ui = g[2], uip1 = g[3], uip2 = g[4];
    const double beta0 = (13. / 12.) * pow(uim2 - 2 * uim1 + ui, 2) + 0.25 * pow(uim2 - 4 * uim1 + 3 * ui, 2);
    const double beta1 = (13. / 12.) * pow(uim1 - 2 * ui + uip1, 2) + 0.25 * pow(uim1 - uip1, 2);
    const double beta2 = (13. / 12.) * pow(ui - 2 * uip1 + uip2, 2) + 0.25 * pow(3 * ui - 4 * uip1 + uip2, 2);
    const double u1 = (1. / 3.) * uim2 - (7. / 6.) * uim1 + (11. / 6.) * ui;
    const double u2 = (-1. / 6.) * uim1 + (5. / 6.) * ui + (1. / 3.) * uip1;
    const double u3 = (1. / 3.) * ui + (5. / 6.) * uip1 - (1. / 6.) * uip2;
    const double omt0 = 0.1 * pow(eps + beta0, -pw);
    const double omt1 = 0.6 * pow(eps + beta1, -pw);
    const double omt2 = 0.3 * pow(eps + beta2, -pw);
    const double omts = omt0 + omt1 + omt2;
    return (omt0 * u1 + omt1 * u2 + omt2 * u3) / omts;
}

void orc_weno5_js_weights(double eps, double pw, const double* g, double* w)
{
    const double uim2 = g[0], uim1 = g[1], ui = g[2], uip1 = g[3], uip2 = g[4];
    const double beta0 = (13. / 12.) * pow(uim2 - 2 * uim1 + ui, 2) + 0.25 * pow(uim2 - 4 * uim1 + 3 * ui, 2);
    const double beta1 = (13. / 12.) * pow(uim1 - 2 * ui + uip1, 2) + 0.25 * pow(uim1 - uip1, 2);
    const double beta2 = (13. / 12.) * pow(ui - 2 * uip1 + uip2, 2) + 0.25 * pow(3 * ui - 4 * uip1 + uip2, 2);
    const double omt0 = 0.1 * pow(eps + beta0, -pw);
    const double omt1 = 0.6 * pow(eps + beta1, -pw);
    const double omt2 = 0.3 * pow(eps + beta2, -pw);
    const double omts = omt0 + omt1 + omt2;
    w[0] = omt0 / omts;
    w[1] = omt1 / omts;
    w[2] = omt2 / omts;
}

/* WenoReconstruct_Z5, lib/WenoReconstruct.cpp:385-438 */
static double weno_z5(double eps, double pw, const double* g)
{
    const double uim2 = g[0], uim1 = g[1], ui = g[2], uip1 = g[3], uip2 = g[4];
    const double beta0 = (13. / 12.) * pow(uim2 - 2 * uim1 + ui, 2) + 0.25 * pow(uim2 - 4 * uim1 + 3 * ui, 2);
    const double beta1 = (13. / 12.) * pow(uim1 - 2 * ui + uip1, 2) + 0.25 * pow(uim1 - uip1, 2);
    const double beta2 = (13. / 12.) * pow(ui - 2 * uip1 + uip2, 2) + 0.25 * pow(3 * ui - 4 * uip1 + uip2, 2);
    const double tau5 = fabs(beta0 - beta2);
    const double u1 = (1. / 3.) * uim2 - (7. / 6.) * uim1 + (11. / 6.) * ui;
    const double u2 = (-1. / 6.) * uim1 + (5. / 6.) * ui + (1. / 3.) * uip1;
    const double u3 = (1. / 3.) * ui + (5. / 6.) * uip1 - (1. / 6.) * uip2;
    const double omt0 = 0.1 * (1. + pow(tau5 / (eps + beta0), pw));
    const double omt1 = 0.6 * (1. + pow(tau5 / (eps + beta1), pw));
    const double omt2 = 0.3 * (1. + pow(tau5 / (eps + beta2), pw));
    const double omts = omt0 + omt1 + omt2;
    return (omt0 * u1 + omt1 * u2 + omt2 * u3) / omts;
}

/* WenoReconstruct_FD5, lib/WenoReconstruct.cpp:580-616 */
static double weno_fd5(const double* g)
{
    const double uim2 = g[0], uim1 = g[1], ui = g[2], uip1 = g[3], uip2 = g[4];
    const double u1 = (1. / 3.) * uim2 - (7. / 6.) * uim1 + (11. / 6.) * ui;
    const double u2 = (-1. / 6.) * uim1 + (5. / 6.) * ui + (1. / 3.) * uip1;
    const double u3 = (1. / 3.) * ui + (5. / 6.) * uip1 - (1. / 6.) * uip2;
    return 0.1 * u1 + 0.6 * u2 + 0.3 * u3;
}

double orc_weno5(int version, double eps, double pw, const double* g)
{
    switch (version) {
    case FINESS_WENO_FD: return weno_fd5(g);
    case FINESS_WENO_Z: return weno_z5(eps, pw, g);
    default: return weno_js5(eps, pw, g);
    }
}

/* Wider reconstructions: Jiang-Shu orders 7, 9, 11 (lib/WenoReconstruct.cpp:
 * 150-197, 199-279, 281-375), WENO-Z 7 and 9 (:440-495, :497-572), linear
 * weights 7 and 9 (:618-656, :658-700) and the first-order value
 * (WenoReconstructLLF, :702-713).
 *
 * Table form, r = (order + 1) / 2, stencil g[0..2r-2] with g[r-1] = u_i:
 * candidate k uses g[k..k+r-1];
 *   u_k    = sum_j CAND[k][j] g[k+j]               (value at x_{i+1/2})
 *   beta_k = sum_{a<=b} B_k[a][b] g[k+a] g[k+b]    (smoothness indicator)
 *   D[k]   linear weights.
 * The indicators carry the reference's scaling: 240 (r = 4), 5040 (r = 5)
 * and 120960 (r = 6) times sum_l int_cell dx^(2l-1) (p_k^(l))^2.  Only the
 * upper triangles of stencils k <= (r-1)/2 are listed (row by row); stencil
 * r-1-k is the mirror image, B_{r-1-k}[a][b] = B_k[r-1-b][r-1-a]. */
static const double CAND4[4][4] = {{-1. / 4., 13. / 12., -23. / 12., 25. / 12.},
                                   {1. / 12., -5. / 12., 13. / 12., 1. / 4.},
                                   {-1. / 12., 7. / 12., 7. / 12., -1. / 12.},
                                   {1. / 4., 13. / 12., -5. / 12., 1. / 12.}};
static const double D4[4] = {1. / 35., 12. / 35., 18. / 35., 4. / 35.};
static const double BH4[2][10] = {
    {547, -3882, 4642, -1854, 7043, -17246, 7042, 11003, -9402, 2107},
    {267, -1642, 1602, -494, 2843, -5966, 1922, 3443, -2522, 547}};

static const double CAND5[5][5] = {{1. / 5., -21. / 20., 137. / 60., -163. / 60., 137. / 60.},
                                   {-1. / 20., 17. / 60., -43. / 60., 77. / 60., 1. / 5.},
                                   {1. / 30., -13. / 60., 47. / 60., 9. / 20., -1. / 20.},
                                   {-1. / 20., 9. / 20., 47. / 60., -13. / 60., 1. / 30.},
                                   {1. / 5., 77. / 60., -43. / 60., 17. / 60., -1. / 20.}};
static const double D5[5] = {1. / 126., 10. / 63., 10. / 21., 20. / 63., 5. / 126.};
static const double BH5[3][15] = {
    {22658, -208501, 364863, -288007, 86329, 482963, -1704396, 1358458, -411487, 1521393, -2462076, 758823,
     1020563, -649501, 107918},
    {6908, -60871, 99213, -70237, 18079, 138563, -464976, 337018, -88297, 406293, -611976, 165153, 242723,
     -140251, 22658},
    {6908, -51001, 67923, -38947, 8209, 104963, -299076, 179098, -38947, 231153, -299076, 67923, 104963, -51001,
     6908}};

static const double CAND6[6][6] = {{-1. / 6., 31. / 30., -163. / 60., 79. / 20., -71. / 20., 49. / 20.},
                                   {1. / 30., -13. / 60., 37. / 60., -21. / 20., 29. / 20., 1. / 6.},
                                   {-1. / 60., 7. / 60., -23. / 60., 19. / 20., 11. / 30., -1. / 30.},
                                   {1. / 60., -2. / 15., 37. / 60., 37. / 60., -2. / 15., 1. / 60.},
                                   {-1. / 30., 11. / 30., 19. / 20., -23. / 60., 7. / 60., -1. / 60.},
                                   {1. / 6., 29. / 20., -21. / 20., 37. / 60., -13. / 60., 1. / 30.}};
static const double D6[6] = {1. / 462., 5. / 77., 25. / 77., 100. / 231., 25. / 154., 1. / 77.};
static const double BH6[3][21] = {
    {1152561, -12950184, 29442256, -33918804, 19834350, -4712740, 36480687, -166461044, 192596472, -113206788,
     27060170, 190757572, -444003904, 262901672, -63394124, 260445372, -311771244, 76206736, 94851237, -47460464,
     6150211},
    {271779, -3015728, 6694608, -7408908, 4067018, -880548, 8449957, -37913324, 42405032, -23510468, 5134574,
     43093692, -97838784, 55053752, -12183636, 56662212, -65224244, 14742480, 19365967, -9117992, 1152561},
    {139633, -1429976, 2863984, -2792660, 1325006, -245620, 3824847, -15880404, 15929912, -7727988, 1458762,
     17195652, -35817664, 17905032, -3462252, 19510972, -20427884, 4086352, 5653317, -2380800, 271779}};

/* The tables of order 7, 9 or 11: cand[r][r], beta[r][r][r] (full upper
 * triangles, mirrored stencils filled in, lower triangles zero), d[r].
 * Returns r, or 0 for an order without a table. */
int orc_weno_tables(int order, double* cand, double* beta, double* d)
{
    const int r = (order + 1) / 2;
    const double *C, *D, *BH;
    int nh;
    switch (order) {
    case 7: C = &CAND4[0][0]; D = D4; BH = &BH4[0][0]; nh = 2; break;
    case 9: C = &CAND5[0][0]; D = D5; BH = &BH5[0][0]; nh = 3; break;
    case 11: C = &CAND6[0][0]; D = D6; BH = &BH6[0][0]; nh = 3; break;
    default: return 0;
    }
    const int tri = r * (r + 1) / 2;
    for (int k = 0; k < r; k++) {
        d[k] = D[k];
        for (int j = 0; j < r; j++) cand[k * r + j] = C[k * r + j];
        for (int a = 0; a < r; a++)
            for (int b = 0; b < r; b++) beta[(k * r + a) * r + b] = 0.0;
    }
    for (int k = 0; k < nh; k++) {
        int t = 0;
        for (int a = 0; a < r; a++)
            for (int b = a; b < r; b++, t++) {
                const double v = BH[k * tri + t];
                beta[(k * r + a) * r + b] = v;
                /* mirror: stencil r-1-k, entry (r-1-b, r-1-a) */
                beta[((r - 1 - k) * r + (r - 1 - b)) * r + (r - 1 - a)] = v;
            }
    }
    return r;
}

typedef struct wide_tab {
    int r;
    double cand[6][6], beta[6][6][6], d[6];
} wide_tab;

static const wide_tab* wide_table(int order)
{
    static wide_tab tabs[3];
    static int ready = 0;
#pragma omp critical(orc_wide_tables)
    if (!ready) {
        for (int t = 0; t < 3; t++) {
            const int o = 7 + 2 * t, r = (o + 1) / 2;
            double c[36], b[216], dd[6];
            tabs[t].r = orc_weno_tables(o, c, b, dd);
            for (int k = 0; k < r; k++) {
                tabs[t].d[k] = dd[k];
                for (int j = 0; j < r; j++) tabs[t].cand[k][j] = c[k * r + j];
                for (int a = 0; a < r; a++)
                    for (int bb = 0; bb < r; bb++) tabs[t].beta[k][a][bb] = b[(k * r + a) * r + bb];
            }
        }
        ready = 1;
    }
    return (order == 7 || order == 9 || order == 11) ? &tabs[(order - 7) / 2] : NULL;
}

/* WENO of order 7, 9 or 11: JS weights d_k (eps + beta_k)^-p; Z weights
 * d_k (1 + (tau / (eps + beta_k))^p) with tau = |beta_0 + 3 (beta_1 - beta_2)
 * - beta_3| (order 7, :470) or |beta_4 - beta_0| (order 9, :549; the
 * reference's std::abs of a double); FD the linear weights. */
static double weno_wide(int version, int order, double eps, double pw, const double* g)
{
    const wide_tab* T = wide_table(order);
    const int r = T->r;
    double u[6], beta[6];
    for (int k = 0; k < r; k++) {
        double s = 0.0, b = 0.0;
        for (int j = 0; j < r; j++) s += T->cand[k][j] * g[k + j];
        for (int a = 0; a < r; a++)
            for (int c = a; c < r; c++) b += T->beta[k][a][c] * g[k + a] * g[k + c];
        u[k] = s;
        beta[k] = b;
    }
    double num = 0.0, den = 0.0;
    if (version == FINESS_WENO_FD) {
        for (int k = 0; k < r; k++) num += T->d[k] * u[k];
        return num;
    }
    double tau = 0.0;
    if (version == FINESS_WENO_Z)
        tau = order == 7 ? fabs(beta[0] + 3. * (beta[1] - beta[2]) - beta[3]) : fabs(beta[4] - beta[0]);
    for (int k = 0; k < r; k++) {
        const double omt = version == FINESS_WENO_Z ? T->d[k] * (1. + pow(tau / (eps + beta[k]), pw))
                                                    : T->d[k] * pow(eps + beta[k], -pw);
        num += omt * u[k];
        den += omt;
    }
    return num / den;
}

/* Reconstruction of any implemented order (GetWenoReconstruct,
 * lib/WenoReconstruct.cpp:39-73): g[0..order-1]. */
double orc_weno(int version, int order, double eps, double pw, const double* g)
{
    if (order == 1) return g[0];
    if (order == 5) return orc_weno5(version, eps, pw, g);
    return weno_wide(version, order, eps, pw, g);
}

/* Whether GetWenoReconstruct has the (version, order) pair. */
int orc_weno_supported(int version, int order)
{
    if (order == 1 || order == 5) return 1;
    if (version == FINESS_WENO_JS) return order == 7 || order == 9 || order == 11;
    if (version == FINESS_WENO_Z || version == FINESS_WENO_FD) return order == 7 || order == 9;
    return 0;
}

/* ====================================================================== */
/* physics callbacks                                                        */
/* ====================================================================== */

/* FluxFunc: Euler apps/2d/euler/lib/FluxFunc.cpp:9-42, shallow water
 * apps/2d/shallow_water/lib/FluxFunc.cpp:8-35, advection
 * apps/1d/advection/lib/FluxFunc.cpp (f = q*aux(1)). */
void orc_flux(const finess_params* p, const double* Q, const double* aux, double* F, double* G)
{
    if (p->physics == FINESS_PHYS_EULER) {
        const double gamma = p->gamma;
        const double rho = Q[0];
        const double u1 = Q[1] / rho, u2 = Q[2] / rho, u3 = Q[3] / rho;
        const double energy = Q[4];
        const double press = (gamma - 1.0e0) * (energy - 0.5e0 * rho * (u1 * u1 + u2 * u2 + u3 * u3));
        F[0] = rho * u1;
        F[1] = rho * u1 * u1 + press;
        F[2] = rho * u1 * u2;
        F[3] = rho * u1 * u3;
        F[4] = u1 * (energy + press);
        if (G) {
            G[0] = rho * u2;
            G[1] = rho * u2 * u1;
            G[2] = rho * u2 * u2 + press;
            G[3] = rho * u2 * u3;
            G[4] = u2 * (energy + press);
        }
    } else if (p->physics == FINESS_PHYS_SHALLOW_WATER) {
        const double h = Q[0], u1 = Q[1] / h, u2 = Q[2] / h;
        F[0] = h * u1;
        F[1] = h * u1 * u1 + 0.5 * h * h;
        F[2] = h * u1 * u2;
        if (G) {
            G[0] = h * u2;
            G[1] = h * u2 * u1;
            G[2] = h * u2 * u2 + 0.5 * h * h;
        }
    } else { /* advection */
        F[0] = Q[0] * aux[0];
        if (G) G[0] = 0.0;
    }
}

/* ProjectLeftEig: Euler apps/2d/euler/lib/ProjectLeftEig.cpp:9-67,
 * shallow water apps/2d/shallow_water/lib/ProjectLeftEig.cpp,
 * advection apps/1d/advection/lib/ProjectLeftEig.cpp (identity). */
void orc_project_left(const finess_params* p, int ixy, const double* Qa, int npts,
                      const double* Qv, double* Wv)
{
#define QV(m, k) Qv[(size_t)((m)-1) * npts + (k)-1]
#define WV(m, k) Wv[(size_t)((m)-1) * npts + (k)-1]
    if (p->physics == FINESS_PHYS_EULER) {
        const int mu = ixy == 1 ? 2 : 3, mv = ixy == 1 ? 3 : 2, mw = 4;
        const double gamma = p->gamma;
        const double rho = Qa[0];
        const double u1 = Qa[mu - 1] / rho, u2 = Qa[mv - 1] / rho, u3 = Qa[mw - 1] / rho;
        const double energy = Qa[4];
        const double umag2 = (u1 * u1 + u2 * u2 + u3 * u3);
        const double press = (gamma - 1.0e0) * (energy - 0.5e0 * rho * umag2);
        const double c = sqrt(gamma * press / rho);
        const double H = (energy + press) / rho;
        for (int k = 1; k <= npts; k++) {
            WV(1, k) = ((umag2 / 2.0 - H - u1 * c) * QV(mu, k) + (umag2 / 2.0 * (c - u1) + H * u1) * QV(1, k)
                        + c * (QV(5, k) - u2 * QV(mv, k) - u3 * QV(mw, k))) / (c * (2.0 * H - umag2));
            WV(mu, k) = 2.0 * ((H - umag2) * QV(1, k) + u1 * QV(mu, k) + u2 * QV(mv, k) + u3 * QV(mw, k) - QV(5, k))
                        / (2.0 * H - umag2);
            WV(mv, k) = QV(mv, k) - u2 * QV(1, k);
            WV(mw, k) = QV(mw, k) - u3 * QV(1, k);
            WV(5, k) = ((H - umag2 / 2.0 - u1 * c) * QV(mu, k) + (umag2 / 2.0 * (c + u1) - H * u1) * QV(1, k)
                        + c * (QV(5, k) - u2 * QV(mv, k) - u3 * QV(mw, k))) / (c * (2.0 * H - umag2));
        }
    } else if (p->physics == FINESS_PHYS_SHALLOW_WATER) {
        const int mu = ixy == 1 ? 2 : 3, mv = ixy == 1 ? 3 : 2;
        const double h = Qa[0], u1 = Qa[mu - 1] / h, u2 = Qa[mv - 1] / h;
        for (int k = 1; k <= npts; k++) {
            WV(1, k) = (sqrt(h) + u1) / (2.0 * sqrt(h)) * QV(1, k) - (0.5 / sqrt(h)) * QV(mu, k);
            WV(2, k) = -u2 * QV(1, k) + QV(mv, k);
            WV(3, k) = (sqrt(h) - u1) / (2.0 * sqrt(h)) * QV(1, k) + (0.5 / sqrt(h)) * QV(mu, k);
        }
    } else {
        for (int k = 1; k <= npts; k++) WV(1, k) = QV(1, k);
    }
#undef QV
#undef WV
}

/* ProjectRightEig: Euler apps/2d/euler/lib/ProjectRightEig.cpp:9-63,
 * shallow water apps/2d/shallow_water/lib/ProjectRightEig.cpp. */
void orc_project_right(const finess_params* p, int ixy, const double* Qa, int npts,
                       const double* Wv, double* Qv)
{
#define QV(m, k) Qv[(size_t)((m)-1) * npts + (k)-1]
#define WV(m, k) Wv[(size_t)((m)-1) * npts + (k)-1]
    if (p->physics == FINESS_PHYS_EULER) {
        const int mu = ixy == 1 ? 2 : 3, mv = ixy == 1 ? 3 : 2, mw = 4;
        const double gamma = p->gamma;
        const double rho = Qa[0];
        const double u1 = Qa[mu - 1] / rho, u2 = Qa[mv - 1] / rho, u3 = Qa[mw - 1] / rho;
        const double energy = Qa[4];
        const double umag2 = (u1 * u1 + u2 * u2 + u3 * u3);
        const double press = (gamma - 1.0e0) * (energy - 0.5e0 * rho * umag2);
        const double c = sqrt(fabs(gamma * press / rho));
        const double H = (energy + press) / rho;
        for (int k = 1; k <= npts; k++) {
            QV(1, k) = WV(1, k) + WV(mu, k) + WV(5, k);
            QV(mu, k) = (u1 - c) * WV(1, k) + u1 * WV(mu, k) + (u1 + c) * WV(5, k);
            QV(mv, k) = u2 * (WV(1, k) + WV(mu, k) + WV(5, k)) + WV(mv, k);
            QV(mw, k) = u3 * (WV(1, k) + WV(mu, k) + WV(5, k)) + WV(mw, k);
            QV(5, k) = (H - u1 * c) * WV(1, k) + umag2 / 2.0 * WV(mu, k) + u2 * WV(mv, k) + u3 * WV(mw, k)
                       + (H + u1 * c) * WV(5, k);
        }
    } else if (p->physics == FINESS_PHYS_SHALLOW_WATER) {
        const int mu = ixy == 1 ? 2 : 3, mv = ixy == 1 ? 3 : 2;
        const double h = Qa[0], u1 = Qa[mu - 1] / h, u2 = Qa[mv - 1] / h;
        for (int k = 1; k <= npts; k++) {
            QV(1, k) = WV(1, k) + WV(3, k);
            QV(mu, k) = (u1 - sqrt(h)) * WV(1, k) + (u1 + sqrt(h)) * WV(3, k);
            QV(mv, k) = u2 * WV(1, k) + WV(2, k) + u2 * WV(3, k);
        }
    } else {
        for (int k = 1; k <= npts; k++) QV(1, k) = WV(1, k);
    }
#undef QV
#undef WV
}

/* SetWaveSpd: Euler apps/2d/euler/lib/SetWaveSpd.cpp:11-69, shallow water
 * apps/2d/shallow_water/lib/SetWaveSpd.cpp:9-48 (u_r from h_l when
 * sw_compat_hl, as at :27-28), advection apps/1d/advection/lib/SetWaveSpd.cpp. */
void orc_wave_speed(const finess_params* p, int ixy, const double* Ql, const double* Qr,
                    const double* auxl, const double* auxr, double* s1, double* s2)
{
    const double n1 = ixy == 1 ? 1.0 : 0.0, n2 = ixy == 1 ? 0.0 : 1.0;
    if (p->physics == FINESS_PHYS_EULER) {
        const double gamma = p->gamma;
        const double rhol = Ql[0];
        const double u1l = Ql[1] / rhol, u2l = Ql[2] / rhol, u3l = Ql[3] / rhol;
        const double energyl = Ql[4];
        const double pressl = (gamma - 1.0e0) * (energyl - 0.5e0 * rhol * (u1l * u1l + u2l * u2l + u3l * u3l));
        const double rhor = Qr[0];
        const double u1r = Qr[1] / rhor, u2r = Qr[2] / rhor, u3r = Qr[3] / rhor;
        const double energyr = Qr[4];
        const double pressr = (gamma - 1.0e0) * (energyr - 0.5e0 * rhor * (u1r * u1r + u2r * u2r + u3r * u3r));
        const double rho = 0.5e0 * (rhol + rhor);
        const double u1 = 0.5e0 * (u1l + u1r);
        const double u2 = 0.5e0 * (u2l + u2r);
        const double press = 0.5e0 * (pressl + pressr);
        const double nmag = sqrt(pow(n1, 2) + pow(n2, 2));
        const double cl = nmag * sqrt(fabs(gamma * pressl / rhol));
        const double cr = nmag * sqrt(fabs(gamma * pressr / rhor));
        const double c = nmag * sqrt(fabs(gamma * press / rho));
        const double un = n1 * u1 + n2 * u2;
        const double unl = n1 * u1l + n2 * u2l;
        const double unr = n1 * u1r + n2 * u2r;
        *s1 = Min(unl - cl, un - c);
        *s2 = Max(unr + cr, un + c);
    } else if (p->physics == FINESS_PHYS_SHALLOW_WATER) {
        const double hl = Ql[0], u1l = Ql[1] / hl, u2l = Ql[2] / hl;
        const double unl = u1l * n1 + u2l * n2;
        const double hr = Qr[0];
        const double hden = p->sw_compat_hl ? hl : hr;
        const double u1r = Qr[1] / hden, u2r = Qr[2] / hden;
        const double unr = u1r * n1 + u2r * n2;
        const double h = 0.5 * (hl + hr);
        const double u1 = (sqrt(hl) * u1l + sqrt(hr) * u1r) / (sqrt(hl) + sqrt(hr));
        const double u2 = (sqrt(hl) * u2l + sqrt(hr) * u2r) / (sqrt(hl) + sqrt(hr));
        const double un = u1 * n1 + u2 * n2;
        *s1 = Min(un - sqrt(h), unl - sqrt(hl));
        *s2 = Max(un + sqrt(h), unr + sqrt(hr));
    } else {
        (void)Ql;
        (void)Qr;
        *s1 = Min(auxl[0], 0.5e0 * (auxl[0] + auxr[0]));
        *s2 = Max(auxr[0], 0.5e0 * (auxl[0] + auxr[0]));
    }
}

/* ---- hyper-dual numbers: exact first and second directional derivatives */
typedef struct { double v, a, b, ab; } hd;
static inline hd hd_c(double x) { hd r = {x, 0, 0, 0}; return r; }
static inline hd hd_add(hd x, hd y) { hd r = {x.v + y.v, x.a + y.a, x.b + y.b, x.ab + y.ab}; return r; }
static inline hd hd_sub(hd x, hd y) { hd r = {x.v - y.v, x.a - y.a, x.b - y.b, x.ab - y.ab}; return r; }
static inline hd hd_scale(double s, hd x) { hd r = {s * x.v, s * x.a, s * x.b, s * x.ab}; return r; }
static inline hd hd_mul(hd x, hd y)
{
    hd r = {x.v * y.v, x.v * y.a + x.a * y.v, x.v * y.b + x.b * y.v,
            x.v * y.ab + x.a * y.b + x.b * y.a + x.ab * y.v};
    return r;
}
static inline hd hd_inv(hd x)
{
    const double i = 1.0 / x.v, i2 = i * i, i3 = i2 * i;
    hd r = {i, -x.a * i2, -x.b * i2, -x.ab * i2 + 2.0 * x.a * x.b * i3};
    return r;
}
static inline hd hd_div(hd x, hd y) { return hd_mul(x, hd_inv(y)); }

/* Flux of the configured physics in hyper-dual arithmetic (same expressions
 * as orc_flux). */
static void flux_hd(const finess_params* p, const hd* Q, hd* F, hd* G)
{
    if (p->physics == FINESS_PHYS_EULER) {
        const double gm1 = p->gamma - 1.0;
        const hd rho = Q[0];
        const hd u1 = hd_div(Q[1], rho), u2 = hd_div(Q[2], rho), u3 = hd_div(Q[3], rho);
        const hd e = Q[4];
        const hd ke = hd_add(hd_add(hd_mul(u1, u1), hd_mul(u2, u2)), hd_mul(u3, u3));
        const hd press = hd_scale(gm1, hd_sub(e, hd_scale(0.5, hd_mul(rho, ke))));
        const hd ru1 = hd_mul(rho, u1), ru2 = hd_mul(rho, u2), ep = hd_add(e, press);
        F[0] = ru1;
        F[1] = hd_add(hd_mul(ru1, u1), press);
        F[2] = hd_mul(ru1, u2);
        F[3] = hd_mul(ru1, u3);
        F[4] = hd_mul(u1, ep);
        G[0] = ru2;
        G[1] = hd_mul(ru2, u1);
        G[2] = hd_add(hd_mul(ru2, u2), press);
        G[3] = hd_mul(ru2, u3);
        G[4] = hd_mul(u2, ep);
    } else if (p->physics == FINESS_PHYS_SHALLOW_WATER) {
        const hd h = Q[0], u1 = hd_div(Q[1], h), u2 = hd_div(Q[2], h);
        const hd hh = hd_scale(0.5, hd_mul(h, h));
        const hd hu1 = hd_mul(h, u1), hu2 = hd_mul(h, u2);
        F[0] = hu1;
        F[1] = hd_add(hd_mul(hu1, u1), hh);
        F[2] = hd_mul(hu1, u2);
        G[0] = hu2;
        G[1] = hd_mul(hu2, u1);
        G[2] = hd_add(hd_mul(hu2, u2), hh);
    } else {
        F[0] = Q[0];
        G[0] = hd_c(0.0);
    }
}

/* DFluxFunc: A[d][m1][m2] = d flux_d(m1) / d q(m2). */
void orc_dflux(const finess_params* p, const double* Q, double* A)
{
    const int meqn = p->meqn;
    for (int m2 = 0; m2 < meqn; m2++) {
        hd q[MAXEQ], F[MAXEQ], G[MAXEQ];
        for (int m = 0; m < meqn; m++) q[m] = hd_c(Q[m]);
        q[m2].a = 1.0;
        flux_hd(p, q, F, G);
        for (int m1 = 0; m1 < meqn; m1++) {
            A[(0 * meqn + m1) * meqn + m2] = F[m1].a;
            A[(1 * meqn + m1) * meqn + m2] = G[m1].a;
        }
    }
}

/* D2FluxFunc: H[d][m][m1][m2] = d^2 flux_d(m) / d q(m1) d q(m2). */
void orc_d2flux(const finess_params* p, const double* Q, double* H)
{
    const int meqn = p->meqn;
    for (int m1 = 0; m1 < meqn; m1++)
        for (int m2 = 0; m2 < meqn; m2++) {
            hd q[MAXEQ], F[MAXEQ], G[MAXEQ];
            for (int m = 0; m < meqn; m++) q[m] = hd_c(Q[m]);
            q[m1].a += 1.0;
            q[m2].b += 1.0;
            flux_hd(p, q, F, G);
            for (int m = 0; m < meqn; m++) {
                H[((0 * meqn + m) * meqn + m1) * meqn + m2] = F[m].ab;
                H[((1 * meqn + m) * meqn + m1) * meqn + m2] = G[m].ab;
            }
        }
}

/* ====================================================================== */
/* boundary conditions                                                      */
/* ====================================================================== */

/* Source index of a ghost index g along an axis of n interior cells. */
static int bc_src(int bc, int g, int n, int* flip)
{
    *flip = 0;
    if (g >= 1 && g <= n) return g;
    if (bc == FINESS_BC_PERIODIC) return g < 1 ? g + n : g - n;
    if (bc == FINESS_BC_REFLECT) {
        *flip = 1;
        return g < 1 ? 1 - g : 2 * n + 1 - g;
    }
    return g < 1 ? 1 : n; /* zeroth-order extrapolation */
}

/* SetBndValues.  2D: sides as apps/2d/euler/riemann_test/SetBndValues.cpp:9-114
 * (extrapolation; also copies q(.,m) into aux(.,m), :37-44) and
 * apps/2d/advection/smooth_example/SetBndValues.cpp:28-194 (periodic with
 * corners).  Corners are filled by applying the x rule then the y rule.
 * 1D: apps/1d/advection/smooth_example/SetBndValues.cpp:6-63. */
void orc_set_bnd(const finess_params* p, double* q, double* aux)
{
    const int mx = p->mx, mbc = p->mbc, meqn = p->meqn, na = naux(p);
    if (p->ndims == 1) {
        for (int i = 1 - mbc; i <= mx + mbc; i++) {
            if (i >= 1 && i <= mx) continue;
            int fl;
            const int bcs = i < 1 ? p->bc[0] : p->bc[1];
            const int si = bc_src(bcs, i, mx, &fl);
            for (int m = 1; m <= meqn; m++) q[ix2(i, m, meqn, mbc)] = q[ix2(si, m, meqn, mbc)];
            if (aux && bcs == FINESS_BC_PERIODIC)
                for (int m = 1; m <= na; m++) aux[ix2(i, m, na, mbc)] = aux[ix2(si, m, na, mbc)];
        }
        return;
    }
    const int my = p->my;
    const int mu_x = 2, mu_y = 3; /* momentum component normal to x / y walls */
    for (int i = 1 - mbc; i <= mx + mbc; i++)
        for (int j = 1 - mbc; j <= my + mbc; j++) {
            if (i >= 1 && i <= mx && j >= 1 && j <= my) continue;
            int fx, fy;
            const int bx = i < 1 ? p->bc[0] : p->bc[1];
            const int by = j < 1 ? p->bc[2] : p->bc[3];
            const int si = bc_src(bx, i, mx, &fx);
            const int sj = bc_src(by, j, my, &fy);
            for (int m = 1; m <= meqn; m++) {
                double v = q[ix3(si, sj, m, my, meqn, mbc)];
                if (p->physics != FINESS_PHYS_ADVECTION) {
                    if (fx && m == mu_x) v = -v;
                    if (fy && m == mu_y) v = -v;
                }
                q[ix3(i, j, m, my, meqn, mbc)] = v;
            }
            if (aux) {
                const int inside_x = (i >= 1 && i <= mx), inside_y = (j >= 1 && j <= my);
                for (int m = 1; m <= na; m++) {
                    double v;
                    if (bx == FINESS_BC_EXTRAP && !inside_x && inside_y && m <= p->maux)
                        v = q[ix3(si, sj, m, my, meqn, mbc)]; /* riemann_test :37-44 quirk */
                    else
                        v = aux[ix3(si, sj, m, my, na, mbc)];
                    aux[ix3(i, j, m, my, na, mbc)] = v;
                }
            }
        }
}

/* ====================================================================== */
/* ConstructL                                                               */
/* ====================================================================== */

/* The configurations the oracle restates: every order GetWenoReconstruct
 * has for ConstructL and the Runge-Kutta stepper (mbc >= r,
 * lib/2d/ConstructL.cpp:31-33, lib/1d/ConstructL.cpp:32-34), order 5 for
 * everything else (check_config5). */
static int check_config(const finess_params* p)
{
    if (p->weno_version != FINESS_WENO_JS && p->weno_version != FINESS_WENO_FD && p->weno_version != FINESS_WENO_Z)
        return FINESS_ERR_CONFIG;
    if (!orc_weno_supported(p->weno_version, p->space_order)) return FINESS_ERR_CONFIG;
    if (p->mbc < 3 || p->mbc < (p->space_order + 1) / 2) return FINESS_ERR_CONFIG;
    if (p->meqn > MAXEQ) return FINESS_ERR_CONFIG;
    if (p->source_term) {
        if (p->source_kind < FINESS_SRC_ZERO || p->source_kind > FINESS_SRC_DAMPING) return FINESS_ERR_CONFIG;
        if (p->source_kind == FINESS_SRC_SW_BATHYMETRY && (p->physics != FINESS_PHYS_SHALLOW_WATER || p->maux < 3))
            return FINESS_ERR_CONFIG;
    }
    return FINESS_OK;
}

/* SourceTermFunc at one point, in the forms the apps define: psi = 0 (e.g.
 * apps/2d/euler/mach_reflection/SourceTermFunc.cpp), the bathymetry term
 * psi = (0, -h aux(2), -h aux(3)) of apps/2d/shallow_water/riemann_test/
 * SourceTermFunc.cpp, or the relaxation psi = -k q. */
static void source_term_func(const finess_params* p, const double* q, const double* aux, double* psi)
{
    for (int m = 0; m < p->meqn; m++) psi[m] = 0.0;
    if (p->source_kind == FINESS_SRC_DAMPING) {
        for (int m = 0; m < p->meqn; m++) psi[m] = -p->source_coef * q[m];
    } else if (p->source_kind == FINESS_SRC_SW_BATHYMETRY) {
        const double h = q[0], bx = aux[1], by = aux[2];
        psi[1] = -h * bx;
        psi[2] = -h * by;
    }
}

static int check_config5(const finess_params* p)
{
    const int rc = check_config(p);
    return rc ? rc : (p->space_order == 5 ? FINESS_OK : FINESS_ERR_CONFIG);
}

/* One interface: characteristic-wise WENO with local Lax-Friedrichs splitting
 * (lib/2d/ConstructL.cpp:63-208 for ixy=1, :215-358 for ixy=2).
 * qs[s][m], fs[s][m]: the ws+1 stencil states / fluxes (cells i-r..i+r-1 for
 * ws = space_order = 2r-1), Ql = qs[r-1], Qr = qs[r]. Returns alpha; writes
 * fhat[meqn]. */
#define ORC_NS 12
static double interface_flux(const finess_params* p, int ixy, const double (*qs)[MAXEQ],
                             const double (*fs)[MAXEQ], const double* auxl, const double* auxr,
                             double alpha_g, double* fhat)
{
    const int meqn = p->meqn, ws = p->space_order, ns = ws + 1, r = ns / 2;
    double Qavg[MAXEQ];
    for (int m = 0; m < meqn; m++) Qavg[m] = 0.5 * (qs[r][m] + qs[r - 1][m]);
    double qvals[MAXEQ * ORC_NS], fvals[MAXEQ * ORC_NS], wvals[MAXEQ * ORC_NS], gvals[MAXEQ * ORC_NS];
    for (int m = 0; m < meqn; m++)
        for (int s = 0; s < ns; s++) {
            qvals[m * ns + s] = qs[s][m];
            fvals[m * ns + s] = fs[s][m];
        }
    orc_project_left(p, ixy, Qavg, ns, qvals, wvals);
    orc_project_left(p, ixy, Qavg, ns, fvals, gvals);
    double s1, s2;
    orc_wave_speed(p, ixy, qs[r - 1], qs[r], auxl, auxr, &s1, &s2);
    /* alpha_g: the global speed of the direction with global_alpha, else 0
     * (lib/2d/ConstructL.cpp:48-54,172,323) */
    const double alpha = Max(alpha_g, Max(fabs(s1), fabs(s2)));
    const double l_alpha = p->alpha_scaling * alpha;
    double ghat[MAXEQ];
    for (int m = 0; m < meqn; m++) {
        double gp[ORC_NS], gm[ORC_NS];
        for (int s = 1; s <= ws; s++) {
            gp[s - 1] = 0.5 * (gvals[m * ns + s - 1] + l_alpha * wvals[m * ns + s - 1]);
            gm[s - 1] = 0.5 * (gvals[m * ns + ws - s + 1] - l_alpha * wvals[m * ns + ws - s + 1]);
        }
        const double dGp = orc_weno(p->weno_version, ws, p->epsilon, p->power_param, gp);
        const double dGm = orc_weno(p->weno_version, ws, p->epsilon, p->power_param, gm);
        ghat[m] = dGp + dGm;
    }
    orc_project_right(p, ixy, Qavg, 1, ghat, fhat);
    return alpha;
}

/* 2D ConstructL, lib/2d/ConstructL.cpp:13-408.  If Fsrc/Gsrc are non-NULL the
 * stencil fluxes are read from them (ConstructLxWL, lib/2d/ConstructLxWL.cpp:
 * 106-236) instead of FluxFunc. */
static int construct_l_2d_faces(const finess_params* p, const double* q, const double* aux,
                                const double* Fsrc, const double* Gsrc, double* Lstar, double* smax,
                                double* Fhat_out, double* Ghat_out);

static int construct_l_2d(const finess_params* p, const double* q, const double* aux,
                          const double* Fsrc, const double* Gsrc, double* Lstar, double* smax)
{
    return construct_l_2d_faces(p, q, aux, Fsrc, Gsrc, Lstar, smax, NULL, NULL);
}

/* As construct_l_2d; Fhat_out / Ghat_out (if non-NULL, laid out as the
 * reference's Fhat (mx+1, my) and Ghat (mx, my+1) with mbc ghosts) receive the
 * interface fluxes (the pFhat / pGhat outputs of lib/2d/ConstructLxWL.cpp:30-33). */
static int construct_l_2d_faces(const finess_params* p, const double* q, const double* aux,
                                const double* Fsrc, const double* Gsrc, double* Lstar, double* smax,
                                double* Fhat_out, double* Ghat_out)
{
    const int mx = p->mx, my = p->my, meqn = p->meqn, mbc = p->mbc, na = naux(p);
    const double dx = p->dx, dy = p->dy;
    double* Fhat = (double*)calloc((size_t)(mx + 1 + 2 * mbc) * (my + 2 * mbc) * meqn, sizeof(double));
    double* Ghat = (double*)calloc((size_t)(mx + 2 * mbc) * (my + 1 + 2 * mbc) * meqn, sizeof(double));
    if (!Fhat || !Ghat) { free(Fhat); free(Ghat); return FINESS_ERR_ARG; }

    const int ns = p->space_order + 1, r = ns / 2;
    /* GlobalWaveSpd on the state the operator reads (lib/2d/ConstructL.cpp:48-54,
     * lib/2d/ConstructLxWL.cpp:88-96) */
    double alpha1 = 0.0, alpha2 = 0.0;
    if (p->global_alpha) orc_global_wave_spd(p, q, aux, &alpha1, &alpha2);
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= mx + 1; i++)
        for (int j = 1; j <= my; j++) {
            double qs[ORC_NS][MAXEQ], fs[ORC_NS][MAXEQ], gdum[MAXEQ];
            for (int s = 1; s <= ns; s++) {
                const int is = i - 1 + s - r;
                for (int m = 1; m <= meqn; m++) qs[s - 1][m - 1] = q[ix3(is, j, m, my, meqn, mbc)];
                if (Fsrc)
                    for (int m = 1; m <= meqn; m++) fs[s - 1][m - 1] = Fsrc[ix3(is, j, m, my, meqn, mbc)];
                else
                    orc_flux(p, qs[s - 1], &aux[ix3(is, j, 1, my, na, mbc)], fs[s - 1], gdum);
            }
            double fh[MAXEQ];
            const double alpha = interface_flux(p, 1, qs, fs, &aux[ix3(i - 1, j, 1, my, na, mbc)],
                                                &aux[ix3(i, j, 1, my, na, mbc)], alpha1, fh);
            double* sm = &smax[ix3(i, j, 1, my + 1, 2, mbc)];
            *sm = Max(*sm, alpha);
            for (int m = 1; m <= meqn; m++) Fhat[ix3(i, j, m, my, meqn, mbc)] = fh[m - 1];
        }

#pragma omp parallel for schedule(static)
    for (int i = 1; i <= mx; i++)
        for (int j = 1; j <= my + 1; j++) {
            double qs[ORC_NS][MAXEQ], fs[ORC_NS][MAXEQ], fdum[MAXEQ];
            for (int s = 1; s <= ns; s++) {
                const int js = j - 1 + s - r;
                for (int m = 1; m <= meqn; m++) qs[s - 1][m - 1] = q[ix3(i, js, m, my, meqn, mbc)];
                if (Gsrc)
                    for (int m = 1; m <= meqn; m++) fs[s - 1][m - 1] = Gsrc[ix3(i, js, m, my, meqn, mbc)];
                else
                    orc_flux(p, qs[s - 1], &aux[ix3(i, js, 1, my, na, mbc)], fdum, fs[s - 1]);
            }
            double gh[MAXEQ];
            const double alpha = interface_flux(p, 2, qs, fs, &aux[ix3(i, j - 1, 1, my, na, mbc)],
                                                &aux[ix3(i, j, 1, my, na, mbc)], alpha2, gh);
            double* sm = &smax[ix3(i, j, 2, my + 1, 2, mbc)];
            *sm = Max(*sm, alpha);
            for (int m = 1; m <= meqn; m++) Ghat[ix3(i, j, m, my + 1, meqn, mbc)] = gh[m - 1];
        }

    /* Lstar = -(Fhat(i+1)-Fhat(i))/dx - (Ghat(j+1)-Ghat(j))/dy, :388-399; with
     * source_term the sampled psi first, then Lstar = psi + tmp (:371-387).
     * ConstructLxWL exits on a source term (lib/2d/ConstructLxWL.cpp:376-380). */
    if (p->source_term && (Fsrc || Gsrc)) {
        free(Fhat);
        free(Ghat);
        return FINESS_ERR_CONFIG;
    }
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= mx; i++)
        for (int j = 1; j <= my; j++) {
            double psi[MAXEQ];
            if (p->source_term)
                source_term_func(p, &q[ix3(i, j, 1, my, meqn, mbc)], &aux[ix3(i, j, 1, my, na, mbc)], psi);
            for (int m = 1; m <= meqn; m++) {
                double tmp = -(Fhat[ix3(i + 1, j, m, my, meqn, mbc)] - Fhat[ix3(i, j, m, my, meqn, mbc)]) / dx;
                tmp = tmp - (Ghat[ix3(i, j + 1, m, my + 1, meqn, mbc)] - Ghat[ix3(i, j, m, my + 1, meqn, mbc)]) / dy;
                Lstar[ix3(i, j, m, my, meqn, mbc)] = p->source_term ? psi[m - 1] + tmp : tmp;
            }
        }
    if (Fhat_out) memcpy(Fhat_out, Fhat, (size_t)(mx + 1 + 2 * mbc) * (my + 2 * mbc) * meqn * sizeof(double));
    if (Ghat_out) memcpy(Ghat_out, Ghat, (size_t)(mx + 2 * mbc) * (my + 1 + 2 * mbc) * meqn * sizeof(double));
    free(Fhat);
    free(Ghat);
    return FINESS_OK;
}

/* 1D ConstructL, lib/1d/ConstructL.cpp:16-224; smax(i) is overwritten (:149).
 * With Fsrc non-NULL the stencil fluxes are read from it: the 1D ConstructLxWL
 * (lib/1d/ConstructLxWL.cpp:14-237), which exits on a source term (:215-219). */
static int construct_l_1d(const finess_params* p, const double* q, const double* aux, const double* Fsrc,
                          double* Lstar, double* smax)
{
    if (Fsrc && p->source_term) return FINESS_ERR_CONFIG;
    const int mx = p->mx, meqn = p->meqn, mbc = p->mbc, na = naux(p);
    const double dx = p->dx;
    double* fhat = (double*)calloc((size_t)(mx + 1 + 2 * mbc) * meqn, sizeof(double));
    if (!fhat) return FINESS_ERR_ARG;
    const int ns = p->space_order + 1, r = ns / 2;
    /* GlobalWaveSpd, lib/1d/GlobalWaveSpd.cpp: max over the interior of the
     * SetWaveSpd speed with both sides equal to the cell (lib/1d/ConstructL.cpp:45-49) */
    double alpha1 = 0.0;
    if (p->global_alpha)
        for (int i = 1; i <= mx; i++) {
            double s1, s2;
            orc_wave_speed(p, 1, &q[ix2(i, 1, meqn, mbc)], &q[ix2(i, 1, meqn, mbc)], &aux[ix2(i, 1, na, mbc)],
                           &aux[ix2(i, 1, na, mbc)], &s1, &s2);
            alpha1 = Max(alpha1, Max(fabs(s1), fabs(s2)));
        }
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= mx + 1; i++) {
        double qs[ORC_NS][MAXEQ], fs[ORC_NS][MAXEQ];
        for (int s = 1; s <= ns; s++) {
            const int is = i - 1 + s - r;
            for (int m = 1; m <= meqn; m++) qs[s - 1][m - 1] = q[ix2(is, m, meqn, mbc)];
            if (Fsrc)
                for (int m = 1; m <= meqn; m++) fs[s - 1][m - 1] = Fsrc[ix2(is, m, meqn, mbc)];
            else
                orc_flux(p, qs[s - 1], &aux[ix2(is, 1, na, mbc)], fs[s - 1], NULL);
        }
        double fh[MAXEQ];
        const double alpha = interface_flux(p, 1, qs, fs, &aux[ix2(i - 1, 1, na, mbc)],
                                            &aux[ix2(i, 1, na, mbc)], alpha1, fh);
        smax[i + mbc - 1] = alpha;
        for (int m = 1; m <= meqn; m++) fhat[ix2(i, m, meqn, mbc)] = fh[m - 1];
    }
    /* with source_term: Lstar = psi + tmp (lib/1d/ConstructL.cpp:194-205) */
    for (int i = 1; i <= mx; i++) {
        double psi[MAXEQ];
        if (p->source_term) source_term_func(p, &q[ix2(i, 1, meqn, mbc)], &aux[ix2(i, 1, na, mbc)], psi);
        for (int m = 1; m <= meqn; m++) {
            const double tmp = -(fhat[ix2(i + 1, m, meqn, mbc)] - fhat[ix2(i, m, meqn, mbc)]) / dx;
            Lstar[ix2(i, m, meqn, mbc)] = p->source_term ? psi[m - 1] + tmp : tmp;
        }
    }
    free(fhat);
    return FINESS_OK;
}

int orc_construct_l(const finess_params* p, const double* q, const double* aux, double* Lstar, double* smax)
{
    const int rc = check_config(p);
    if (rc) return rc;
    if (p->ndims == 1) return construct_l_1d(p, q, aux, NULL, Lstar, smax);
    return construct_l_2d(p, q, aux, NULL, NULL, Lstar, smax);
}

/* ====================================================================== */
/* CFL, update, steppers                                                    */
/* ====================================================================== */

int orc_get_cfl(const finess_params* p, double dt, double dtmax, const double* smax, double* cfl_out)
{
    double cfl = -100.0;
    if (dt > dtmax) return FINESS_ERR_DT;
    if (p->ndims == 1) { /* lib/1d/GetCFL.cpp:9-45 */
        for (int j = 1; j <= p->mx; j++) cfl = Max(dt * smax[j + p->mbc - 1] / p->dx, cfl);
        *cfl_out = cfl;
        if (cfl > 1.0e10) return FINESS_ERR_CFL;
        return FINESS_OK;
    }
    /* lib/2d/GetCFL.cpp:31-74: i=1..mx, j=1..my only */
    for (int i = 1; i <= p->mx; i++)
        for (int j = 1; j <= p->my; j++) {
            cfl = Max(dt * smax[ix3(i, j, 1, p->my + 1, 2, p->mbc)] / p->dx, cfl);
            cfl = Max(dt * smax[ix3(i, j, 2, p->my + 1, 2, p->mbc)] / p->dy, cfl);
        }
    *cfl_out = cfl;
    if (cfl > 1.0e8) return FINESS_ERR_CFL;
    return FINESS_OK;
}

/* UpdateSoln, lib/2d/UpdateSoln.cpp:14-38: qnew = a1*qstar + a2*qnew + b*dt*L
 * over all numel entries. */
static void update_soln(size_t numel, double a1, double a2, double b, double dt,
                        const double* qstar, const double* L, double* qnew)
{
#pragma omp parallel for schedule(static)
    for (long long k = 0; k < (long long)numel; k++) {
        const double tmp = a1 * qstar[k] + a2 * qnew[k] + b * dt * L[k];
        qnew[k] = tmp;
    }
}

static size_t smax_numel(const finess_params* p)
{
    if (p->ndims == 1) return (size_t)(p->mx + 2 * p->mbc);
    return (size_t)(p->mx + 1 + 2 * p->mbc) * (size_t)(p->my + 1 + 2 * p->mbc) * 2;
}

/* SSP-RK coefficients, lib/2d/SetRKinfo.cpp:34-64 (orders 1-3). */
static const double RK_A1[3][3] = {{1.0, 0, 0}, {1.0, 0.5, 0}, {1.0, 0.75, 2.0e0 / 3.0e0}};
static const double RK_A2[3][3] = {{0.0, 0, 0}, {0.0, 0.5, 0}, {0.0, 0.25, 1.0e0 / 3.0e0}};
static const double RK_B[3][3] = {{1.0, 0, 0}, {1.0, 0.5, 0}, {1.0, 0.25, 2.0e0 / 3.0e0}};

/* The 10-stage fourth-order and the 8-stage fifth-order schemes,
 * lib/2d/SetRKinfo.cpp:66-129 (coefficients as published there). */
static const double RK5_DELTA[8] = {1.0e0, 1.528486658778845e00, 4.720094096662784e-02, 8.801244253465348e-01,
                                    1.019066090228480e+00, 1.049772291176110e+01, -4.254616508506826e+00, 0.0};
static const double RK5_GAMMA[3][8] = {
    {0.0, -1.552288007713033e+01, 4.127286635722417e-01, -1.011819196331377e+00, -2.765748383780848e-01,
     5.075770311217778e-02, 6.999810478513669e+00, -1.114908881433104e+01},
    {1.0, 6.534691420958578e+00, 2.280056542904473e-01, 1.308684311397668e+00, 4.769419552531064e-01,
     -6.368809762042849e-03, 9.339446057238532e-02, 9.556626047962331e-01},
    {0.0, 0.0, 0.0, -2.510747784045939e+00, -8.576822794622042e-01, 1.044599944472272e+00, -7.000810861049136e+00,
     1.906311811144179e+00}};
static const double RK5_BETA[8] = {8.653258038183180e-02, 9.544677980851571e-01, 2.651941386774408e-01,
                                   2.736914413910379e-01, 5.999778649323600e-01, 4.433177471748104e-03,
                                   5.309971130968292e-03, 5.830861806762871e-01};

/* lib/2d/FinSolveRK.cpp:213-291: SSP(10,4).  q1, q2 start as q^n; stages 1-5
 * and 6-9 are q1 = q1 + dt/6 L(q1), between them q2 = (q2 + 9 q1)/25 and
 * q1 = 15 q2 - 5 q1, and the last stage is q1 = q2 + 3/5 q1 + dt/10 L(q1). */
static int rk4_step(const finess_params* p, double* qnew, double* aux, double* L, double* smax, double dt,
                    double* q2)
{
    const size_t numel = state_numel(p, p->meqn);
    int rc;
    memcpy(q2, qnew, numel * sizeof(double));
    for (int s = 1; s <= 9; s++) {
        orc_set_bnd(p, qnew, aux);
        if ((rc = orc_construct_l(p, qnew, aux, L, smax))) return rc;
        update_soln(numel, 1.0, 0.0, 1.0 / 6.0, dt, qnew, L, qnew);
        if (s == 5) {
#pragma omp parallel for schedule(static)
            for (long long k = 0; k < (long long)numel; k++) {
                const double tmp = (q2[k] + 9.0 * qnew[k]) / 25.0;
                q2[k] = tmp;
                qnew[k] = 15.0 * tmp - 5.0 * qnew[k];
            }
        }
    }
    orc_set_bnd(p, qnew, aux);
    if ((rc = orc_construct_l(p, qnew, aux, L, smax))) return rc;
    update_soln(numel, 1.0, 3.0 / 5.0, 1.0 / 10.0, dt, q2, L, qnew);
    return FINESS_OK;
}

/* lib/2d/FinSolveRK.cpp:292-305 with the two-register update of
 * lib/2d/UpdateSoln.cpp:40-80: q2 = 0; for s = 1..8, with q1 = qnew:
 * q2 = q2 + delta_s q1, q1 = g1 q1 + g2 q2 + g3 q^n + beta_s dt L(q1). */
static int rk5_step(const finess_params* p, double* qnew, double* aux, double* L, double* smax, double dt,
                    double* q2, const double* qold)
{
    const size_t numel = state_numel(p, p->meqn);
    int rc;
    memset(q2, 0, numel * sizeof(double));
    for (int s = 0; s < 8; s++) {
        orc_set_bnd(p, qnew, aux);
        if ((rc = orc_construct_l(p, qnew, aux, L, smax))) return rc;
        const double g1 = RK5_GAMMA[0][s], g2 = RK5_GAMMA[1][s], g3 = RK5_GAMMA[2][s], delta = RK5_DELTA[s];
        const double beta = RK5_BETA[s];
#pragma omp parallel for schedule(static)
        for (long long k = 0; k < (long long)numel; k++) {
            const double s1 = qnew[k];
            const double s2 = q2[k] + delta * s1;
            q2[k] = s2;
            const double s3 = qold[k];
            qnew[k] = g1 * s1 + g2 * s2 + g3 * s3 + beta * dt * L[k];
        }
    }
    return FINESS_OK;
}

/* One full RK step (all stages) from qnew into qnew, lib/2d/FinSolveRK.cpp:
 * 115-305 (orders 1-5). qstar, L, smax are scratch. */
static int rk_step(const finess_params* p, double* qnew, double* qstar, double* aux, double* L,
                   double* smax, double dt)
{
    const size_t numel = state_numel(p, p->meqn);
    const int to = p->time_order, k = to - 1;
    int rc;
    if (to == 1) {
        orc_set_bnd(p, qnew, aux);
        if ((rc = orc_construct_l(p, qnew, aux, L, smax))) return rc;
        update_soln(numel, RK_A1[k][0], RK_A2[k][0], RK_B[k][0], dt, qnew, L, qnew);
    } else if (to == 2) {
        orc_set_bnd(p, qnew, aux);
        if ((rc = orc_construct_l(p, qnew, aux, L, smax))) return rc;
        update_soln(numel, RK_A1[k][0], RK_A2[k][0], RK_B[k][0], dt, qnew, L, qstar);
        orc_set_bnd(p, qstar, aux);
        if ((rc = orc_construct_l(p, qstar, aux, L, smax))) return rc;
        update_soln(numel, RK_A1[k][1], RK_A2[k][1], RK_B[k][1], dt, qstar, L, qnew);
    } else if (to == 3) {
        orc_set_bnd(p, qnew, aux);
        if ((rc = orc_construct_l(p, qnew, aux, L, smax))) return rc;
        update_soln(numel, RK_A1[k][0], RK_A2[k][0], RK_B[k][0], dt, qnew, L, qstar);
        orc_set_bnd(p, qstar, aux);
        if ((rc = orc_construct_l(p, qstar, aux, L, smax))) return rc;
        update_soln(numel, RK_A1[k][1], RK_A2[k][1], RK_B[k][1], dt, qnew, L, qstar);
        orc_set_bnd(p, qstar, aux);
        if ((rc = orc_construct_l(p, qstar, aux, L, smax))) return rc;
        update_soln(numel, RK_A1[k][2], RK_A2[k][2], RK_B[k][2], dt, qstar, L, qnew);
    } else if (to == 4 || to == 5) {
        double* q2 = (double*)malloc(numel * sizeof(double));
        double* q0 = (double*)malloc(numel * sizeof(double));
        if (!q2 || !q0) { free(q2); free(q0); return FINESS_ERR_ARG; }
        memcpy(q0, qnew, numel * sizeof(double));
        rc = to == 4 ? rk4_step(p, qnew, aux, L, smax, dt, q2) : rk5_step(p, qnew, aux, L, smax, dt, q2, q0);
        free(q2);
        free(q0);
        return rc;
    } else {
        return FINESS_ERR_CONFIG;
    }
    return FINESS_OK;
}

/* ---- positivity-preserving (MPP) flux limiter of the LxW path ---------- */

/* GlobalWaveSpd, lib/2d/GlobalWaveSpd.cpp:20-80: the largest |s1|, |s2| of
 * SetWaveSpd over the interior cells, each cell against itself. */
void orc_global_wave_spd(const finess_params* p, const double* q, const double* aux, double* alpha1,
                         double* alpha2)
{
    const int na = naux(p);
    double a1 = 0.0, a2 = 0.0;
    for (int i = 1; i <= p->mx; i++)
        for (int j = 1; j <= p->my; j++) {
            const double* Q = &q[ix3(i, j, 1, p->my, p->meqn, p->mbc)];
            const double* A = &aux[ix3(i, j, 1, p->my, na, p->mbc)];
            double s1, s2;
            orc_wave_speed(p, 1, Q, Q, A, A, &s1, &s2);
            a1 = Max(a1, Max(fabs(s1), fabs(s2)));
            orc_wave_speed(p, 2, Q, Q, A, A, &s1, &s2);
            a2 = Max(a2, Max(fabs(s1), fabs(s2)));
        }
    *alpha1 = a1;
    *alpha2 = a2;
}

/* ConstructLFL, lib/2d/ConstructLFL.cpp:25-159: Lax-Friedrichs interface
 * fluxes with the global wave speeds (:112-128); smax(1,1,1) and smax(1,1,2)
 * are set to those speeds (:62-64).  The Lstar it also forms (:140-158) is
 * not used by the limited update and is not restated.  q carries its
 * boundary values. */
static void construct_lfl(const finess_params* p, const double* q, const double* aux, double* fLF,
                          double* gLF, double* smax)
{
    const int mx = p->mx, my = p->my, meqn = p->meqn, mbc = p->mbc, na = naux(p);
    double ax, ay;
    orc_global_wave_spd(p, q, aux, &ax, &ay);
    smax[ix3(1, 1, 1, my + 1, 2, mbc)] = ax;
    smax[ix3(1, 1, 2, my + 1, 2, mbc)] = ay;
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= mx + 1; i++)
        for (int j = 1; j <= my; j++) {
            double fa[MAXEQ], fb[MAXEQ], gd[MAXEQ];
            orc_flux(p, &q[ix3(i - 1, j, 1, my, meqn, mbc)], &aux[ix3(i - 1, j, 1, my, na, mbc)], fa, gd);
            orc_flux(p, &q[ix3(i, j, 1, my, meqn, mbc)], &aux[ix3(i, j, 1, my, na, mbc)], fb, gd);
            for (int m = 1; m <= meqn; m++)
                fLF[ix3(i, j, m, my, meqn, mbc)] = 0.5 * (fa[m - 1] + fb[m - 1]) +
                    0.5 * ax * (q[ix3(i - 1, j, m, my, meqn, mbc)] - q[ix3(i, j, m, my, meqn, mbc)]);
        }
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= mx; i++)
        for (int j = 1; j <= my + 1; j++) {
            double ga[MAXEQ], gb[MAXEQ], fd[MAXEQ];
            orc_flux(p, &q[ix3(i, j - 1, 1, my, meqn, mbc)], &aux[ix3(i, j - 1, 1, my, na, mbc)], fd, ga);
            orc_flux(p, &q[ix3(i, j, 1, my, meqn, mbc)], &aux[ix3(i, j, 1, my, na, mbc)], fd, gb);
            for (int m = 1; m <= meqn; m++)
                gLF[ix3(i, j, m, my + 1, meqn, mbc)] = 0.5 * (ga[m - 1] + gb[m - 1]) +
                    0.5 * ay * (q[ix3(i, j - 1, m, my, meqn, mbc)] - q[ix3(i, j, m, my, meqn, mbc)]);
        }
}

/* Pressure of a 2D Euler state as the limiter forms it. */
static double mpp_pressure(double gm1, const double* Q)
{
    const double rho = Q[0], u1 = Q[1] / rho, u2 = Q[2] / rho, u3 = Q[3] / rho, energy = Q[4];
    return gm1 * (energy - 0.5 * rho * (u1 * u1 + u2 * u2 + u3 * u3));
}

/* ApplyMPPLimiter2D for 2D Euler, apps/2d/euler/lib/ApplyMPPLimiter2D.cpp:
 * 27-286.  Per cell, the largest fractions amin[k] of the four
 * (high - Lax-Friedrichs) flux differences that keep the density above
 * 1e-12 (:56-95) and then the pressure non-negative for every combination
 * of the four faces (16 combinations, 9 bisection steps each, :96-232); each
 * face takes the smaller fraction of its two cells (:236-262) and its flux
 * becomes theta (fHat - fLF) + fLF (:264-284).  fHat, gHat are limited in
 * place. */
static void apply_mpp_limiter_2d(const finess_params* p, double dt, const double* q, const double* fLF,
                                 const double* gLF, double* fHat, double* gHat)
{
    const int mx = p->mx, my = p->my, meqn = p->meqn, mbc = p->mbc;
    const double eps = 1.0e-12, rho_min = eps;
    const double dtx = dt / p->dx, dty = dt / p->dy;
    const double gm1 = p->gamma - 1.0;
    double* amin = (double*)malloc((size_t)mx * my * 4 * sizeof(double));
#define FX(a, b, m) fHat[ix3(a, b, m, my, meqn, mbc)]
#define FL(a, b, m) fLF[ix3(a, b, m, my, meqn, mbc)]
#define GX(a, b, m) gHat[ix3(a, b, m, my + 1, meqn, mbc)]
#define GL(a, b, m) gLF[ix3(a, b, m, my + 1, meqn, mbc)]
#define AM(a, b, k) amin[((size_t)((a) - 1) * my + ((b) - 1)) * 4 + (k)]
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= mx; i++)
        for (int j = 1; j <= my; j++) {
            /* density: the Lax-Friedrichs update's margin above rho_min */
            const double gmin = rho_min - (q[ix3(i, j, 1, my, meqn, mbc)] + dtx * (FL(i, j, 1) - FL(i + 1, j, 1)) +
                                           dty * (GL(i, j, 1) - GL(i, j + 1, 1)));
            double ff[4];
            ff[0] = dtx * (FX(i, j, 1) - FL(i, j, 1));
            ff[1] = -dtx * (FX(i + 1, j, 1) - FL(i + 1, j, 1));
            ff[2] = dty * (GX(i, j, 1) - GL(i, j, 1));
            ff[3] = -dty * (GX(i, j + 1, 1) - GL(i, j + 1, 1));
            int neg[4];
            for (int k = 0; k < 4; k++) neg[k] = ff[k] < 0.0;
            double ffsum = (double)neg[0] * ff[0];
            for (int k = 1; k < 4; k++) ffsum = ffsum + (double)neg[k] * ff[k];
            double a[4];
            for (int k = 0; k < 4; k++) a[k] = neg[k] ? Min(gmin / (ffsum - eps), 1.e0) : 1.e0;

            /* pressure: every on/off combination of the four faces */
            double qlf[MAXEQ], fh[4][MAXEQ], fl[4][MAXEQ], q0[MAXEQ];
            for (int m = 1; m <= meqn; m++) {
                qlf[m - 1] = q[ix3(i, j, m, my, meqn, mbc)] + dtx * (FL(i, j, m) - FL(i + 1, j, m)) +
                             dty * (GL(i, j, m) - GL(i, j + 1, m));
                fh[0][m - 1] = FX(i, j, m);
                fh[1][m - 1] = FX(i + 1, j, m);
                fh[2][m - 1] = GX(i, j, m);
                fh[3][m - 1] = GX(i, j + 1, m);
                fl[0][m - 1] = FL(i, j, m);
                fl[1][m - 1] = FL(i + 1, j, m);
                fl[2][m - 1] = GL(i, j, m);
                fl[3][m - 1] = GL(i, j + 1, m);
                q0[m - 1] = q[ix3(i, j, m, my, meqn, mbc)];
            }
            double rescale[16];
            for (int c = 0; c < 16; c++) {
                /* c = 8 i1 + 4 i2 + 2 i3 + i4: face k is on when bit (3 - k) is set */
                double aa[4];
                for (int k = 0; k < 4; k++) aa[k] = (double)((c >> (3 - k)) & 1) * a[k];
                double qh[MAXEQ];
                for (int m = 0; m < meqn; m++) {
                    double f4[4];
                    for (int k = 0; k < 4; k++) f4[k] = aa[k] * fh[k][m] + (1.0 - aa[k]) * fl[k][m];
                    qh[m] = q0[m] + dtx * (f4[0] - f4[1]) + dty * (f4[2] - f4[3]);
                }
                double phigh = mpp_pressure(gm1, qh);
                rescale[c] = 1.0;
                if (phigh < 0.0) {
                    double root1 = 0.0, root2 = 1.0, rate = 0.0;
                    for (int nn = 1; nn < 10; nn++) {
                        rate = (root1 + root2) / 2.0;
                        double qt[MAXEQ];
                        for (int m = 0; m < meqn; m++) qt[m] = rate * qh[m] + (1.0 - rate) * qlf[m];
                        phigh = mpp_pressure(gm1, qt);
                        if (phigh < 0.0) root2 = rate;
                        else root1 = rate;
                    }
                    if (!(phigh > 0.0)) rate = root1;
                    rescale[c] = rate;
                }
            }
            /* face k: the smallest rescale over the combinations with it on */
            for (int k = 0; k < 4; k++) {
                double r = 1.0;
                for (int c = 0; c < 16; c++)
                    if ((c >> (3 - k)) & 1) r = Min(r, rescale[c]);
                AM(i, j, k) = r * a[k];
            }
        }
    /* theta per face from its two cells, then the limited fluxes */
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= mx + 1; i++)
        for (int j = 1; j <= my; j++) {
            const double th = i == 1 ? AM(1, j, 0) : i == mx + 1 ? AM(mx, j, 1) : Min(AM(i, j, 0), AM(i - 1, j, 1));
            for (int m = 1; m <= meqn; m++) FX(i, j, m) = th * (FX(i, j, m) - FL(i, j, m)) + FL(i, j, m);
        }
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= mx; i++)
        for (int j = 1; j <= my + 1; j++) {
            const double th = j == 1 ? AM(i, 1, 2) : j == my + 1 ? AM(i, my, 3) : Min(AM(i, j, 2), AM(i, j - 1, 3));
            for (int m = 1; m <= meqn; m++) GX(i, j, m) = th * (GX(i, j, m) - GL(i, j, m)) + GL(i, j, m);
        }
#undef FX
#undef FL
#undef GX
#undef GL
#undef AM
    free(amin);
}

static size_t xface_numel(const finess_params* p)
{
    return (size_t)(p->mx + 1 + 2 * p->mbc) * (p->my + 2 * p->mbc) * p->meqn;
}
static size_t yface_numel(const finess_params* p)
{
    return (size_t)(p->mx + 2 * p->mbc) * (p->my + 1 + 2 * p->mbc) * p->meqn;
}

/* The interface fluxes of one limited LxW step (q carries its boundary
 * values): high-order Fh, Gh (ConstructLxWL), Lax-Friedrichs Fl, Gl
 * (ConstructLFL) and the limited Fm, Gm (ApplyMPPLimiter2D); smax as the
 * reference leaves it.  Arrays as construct_l_2d_faces. */
static int lxw_mpp_faces(const finess_params* p, double dt, const double* q, const double* aux, double* F,
                         double* G, double* L, double* smax, double* Fh, double* Gh, double* Fl, double* Gl,
                         double* Fm, double* Gm)
{
    int rc;
    if ((rc = orc_construct_integrated_r(p, dt, q, aux, F, G))) return rc;
    if ((rc = construct_l_2d_faces(p, q, aux, F, G, L, smax, Fm, Gm))) return rc;
    if (Fh) memcpy(Fh, Fm, xface_numel(p) * sizeof(double));
    if (Gh) memcpy(Gh, Gm, yface_numel(p) * sizeof(double));
    construct_lfl(p, q, aux, Fl, Gl, smax);
    apply_mpp_limiter_2d(p, dt, q, Fl, Gl, Fm, Gm);
    return FINESS_OK;
}

int orc_lxw_mpp_faces(const finess_params* p, double dt, const double* q, const double* aux, double* Fh,
                      double* Gh, double* Fl, double* Gl, double* Fm, double* Gm)
{
    int rc = check_config5(p);
    if (rc) return rc;
    const size_t numel = state_numel(p, p->meqn);
    double* F = (double*)calloc(numel, sizeof(double));
    double* G = (double*)calloc(numel, sizeof(double));
    double* L = (double*)calloc(numel, sizeof(double));
    double* smax = (double*)calloc(smax_numel(p), sizeof(double));
    rc = lxw_mpp_faces(p, dt, q, aux, F, G, L, smax, Fh, Gh, Fl, Gl, Fm, Gm);
    free(F); free(G); free(L); free(smax);
    return rc;
}

/* One LxW step, lib/2d/FinSolveLxW.cpp:122-171; with mpp_limiter the
 * limited fluxes update the interior cells, x then y (:126-157). */
static int construct_integrated_f_1d(const finess_params* p, double dt, const double* q, const double* aux,
                                     double* F);

static int lxw_step(const finess_params* p, double* qnew, double* aux, double* F, double* G,
                    double* L, double* smax, double dt)
{
    const size_t numel = state_numel(p, p->meqn);
    int rc;
    if (p->ndims == 1) {
        /* lib/1d/FinSolveLxW.cpp:100-142 without the limiter: SetBndValues,
         * ConstructIntegratedF, ConstructLxWL, q = q + dt L on every entry */
        if (p->mpp_limiter) return FINESS_ERR_CONFIG;
        orc_set_bnd(p, qnew, aux);
        if ((rc = construct_integrated_f_1d(p, dt, qnew, aux, F))) return rc;
        if ((rc = construct_l_1d(p, qnew, aux, F, L, smax))) return rc;
        for (long long k = 0; k < (long long)numel; k++) qnew[k] = qnew[k] + dt * L[k];
        return FINESS_OK;
    }
    orc_set_bnd(p, qnew, aux);
    if ((rc = orc_construct_integrated_r(p, dt, qnew, aux, F, G))) return rc;
    if (p->mpp_limiter) {
        const int mx = p->mx, my = p->my, meqn = p->meqn, mbc = p->mbc;
        double* Fl = (double*)calloc(xface_numel(p), sizeof(double));
        double* Gl = (double*)calloc(yface_numel(p), sizeof(double));
        double* Fm = (double*)calloc(xface_numel(p), sizeof(double));
        double* Gm = (double*)calloc(yface_numel(p), sizeof(double));
        rc = lxw_mpp_faces(p, dt, qnew, aux, F, G, L, smax, NULL, NULL, Fl, Gl, Fm, Gm);
        if (!rc) {
            const double dtdx = dt / p->dx, dtdy = dt / p->dy;
#pragma omp parallel for schedule(static)
            for (int i = 1; i <= mx; i++)
                for (int j = 1; j <= my; j++)
                    for (int m = 1; m <= meqn; m++) {
                        double* Q = &qnew[ix3(i, j, m, my, meqn, mbc)];
                        double tmp = Fm[ix3(i + 1, j, m, my, meqn, mbc)] - Fm[ix3(i, j, m, my, meqn, mbc)];
                        *Q = *Q - dtdx * tmp;
                        tmp = Gm[ix3(i, j + 1, m, my + 1, meqn, mbc)] - Gm[ix3(i, j, m, my + 1, meqn, mbc)];
                        *Q = *Q - dtdy * tmp;
                    }
        }
        free(Fl); free(Gl); free(Fm); free(Gm);
        return rc;
    }
    if ((rc = orc_construct_lxw_l(p, qnew, aux, F, G, L, smax))) return rc;
#pragma omp parallel for schedule(static)
    for (long long k = 0; k < (long long)numel; k++) qnew[k] = qnew[k] + dt * L[k];
    return FINESS_OK;
}

/* ---- FinSolveRK_PIF (lib/2d/FinSolveRK_PIF.cpp) --------------------------- */

/* ConstructL_NOC, lib/2d/FinSolveRK_PIF.cpp:723-977: component-wise WENO of
 * the flux without characteristic projection and without wave-speed
 * splitting (gp = f/2 over cells i-3..i+1, gm = f/2 over i+2..i-2); smax is
 * not touched. */
static int construct_l_noc(const finess_params* p, const double* q, const double* aux, double* Lstar)
{
    const int mx = p->mx, my = p->my, meqn = p->meqn, mbc = p->mbc, na = naux(p), ws = 5;
    const double dx = p->dx, dy = p->dy;
    double* Fhat = (double*)calloc((size_t)(mx + 1 + 2 * mbc) * (my + 2 * mbc) * meqn, sizeof(double));
    double* Ghat = (double*)calloc((size_t)(mx + 2 * mbc) * (my + 1 + 2 * mbc) * meqn, sizeof(double));
    if (!Fhat || !Ghat) { free(Fhat); free(Ghat); return FINESS_ERR_ARG; }
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= mx + 1; i++)
        for (int j = 1; j <= my; j++) {
            double fs[6][MAXEQ], gd[MAXEQ];
            for (int s = 1; s <= ws + 1; s++) {
                const int is = i - 1 + s - 3;
                orc_flux(p, &q[ix3(is, j, 1, my, meqn, mbc)], &aux[ix3(is, j, 1, my, na, mbc)], fs[s - 1], gd);
            }
            for (int m = 0; m < meqn; m++) {
                double gp[5], gm[5];
                for (int s = 1; s <= ws; s++) {
                    gp[s - 1] = 0.5 * fs[s - 1][m];
                    gm[s - 1] = 0.5 * fs[ws - s + 1][m];
                }
                Fhat[ix3(i, j, m + 1, my, meqn, mbc)] = orc_weno5(p->weno_version, p->epsilon, p->power_param, gp) +
                                                        orc_weno5(p->weno_version, p->epsilon, p->power_param, gm);
            }
        }
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= mx; i++)
        for (int j = 1; j <= my + 1; j++) {
            double gs[6][MAXEQ], fd[MAXEQ];
            for (int s = 1; s <= ws + 1; s++) {
                const int js = j - 1 + s - 3;
                orc_flux(p, &q[ix3(i, js, 1, my, meqn, mbc)], &aux[ix3(i, js, 1, my, na, mbc)], fd, gs[s - 1]);
            }
            for (int m = 0; m < meqn; m++) {
                double gp[5], gm[5];
                for (int s = 1; s <= ws; s++) {
                    gp[s - 1] = 0.5 * gs[s - 1][m];
                    gm[s - 1] = 0.5 * gs[ws - s + 1][m];
                }
                Ghat[ix3(i, j, m + 1, my + 1, meqn, mbc)] = orc_weno5(p->weno_version, p->epsilon, p->power_param, gp) +
                                                            orc_weno5(p->weno_version, p->epsilon, p->power_param, gm);
            }
        }
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= mx; i++)
        for (int j = 1; j <= my; j++)
            for (int m = 1; m <= meqn; m++) {
                double tmp = -(Fhat[ix3(i + 1, j, m, my, meqn, mbc)] - Fhat[ix3(i, j, m, my, meqn, mbc)]) / dx;
                tmp = tmp - (Ghat[ix3(i, j + 1, m, my + 1, meqn, mbc)] - Ghat[ix3(i, j, m, my + 1, meqn, mbc)]) / dy;
                Lstar[ix3(i, j, m, my, meqn, mbc)] = tmp;
            }
    free(Fhat);
    free(Ghat);
    return FINESS_OK;
}

/* EulerStep, lib/2d/FinSolveRK_PIF.cpp:980-1004: qnew = qold + dt L on the
 * interior. */
static void euler_step(const finess_params* p, double dt, const double* qold, const double* L, double* qnew)
{
    const int mx = p->mx, my = p->my, meqn = p->meqn, mbc = p->mbc;
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= mx; i++)
        for (int j = 1; j <= my; j++)
            for (int m = 1; m <= meqn; m++) {
                const size_t k = ix3(i, j, m, my, meqn, mbc);
                qnew[k] = qold[k] + dt * L[k];
            }
}

/* Adds w times the flux of q (FluxFunc sampled over every cell, ghosts
 * included, after the boundary values are set) to the running sums Fs, Gs;
 * first: the sums start from 0 (lib/2d/FinSolveRK_PIF.cpp:155-186). */
static void pif_flux_sum(const finess_params* p, const double* q, const double* aux, double w, int first,
                         double* Fs, double* Gs)
{
    const int mx = p->mx, my = p->my, meqn = p->meqn, mbc = p->mbc, na = naux(p);
#pragma omp parallel for schedule(static)
    for (int i = 1 - mbc; i <= mx + mbc; i++)
        for (int j = 1 - mbc; j <= my + mbc; j++) {
            double f[MAXEQ], g[MAXEQ];
            orc_flux(p, &q[ix3(i, j, 1, my, meqn, mbc)], &aux[ix3(i, j, 1, my, na, mbc)], f, g);
            for (int m = 1; m <= meqn; m++) {
                const size_t k = ix3(i, j, m, my, meqn, mbc);
                if (first) { Fs[k] = 0.; Gs[k] = 0.; }
                Fs[k] += w * f[m - 1];
                Gs[k] += w * g[m - 1];
            }
        }
}

/* One RK4-PIF step, lib/2d/FinSolveRK_PIF.cpp:151-206 (time_order 4): three
 * non-characteristic stages to q2, q3, q4, the time-averaged fluxes
 * (R1 + 2 R2 + 2 R3 + R4)/6, then characteristic WENO on them (ConstructL
 * with fstar, gstar, :308-715, the eigen-systems and wave speeds from q^n)
 * and qnew = qold + dt L.  qnew is q^n on entry (its boundary values get
 * set), qstar, Fs, Gs, L are scratch. */
static int pif_step(const finess_params* p, double* qnew, double* qstar, double* aux, double* Fs, double* Gs,
                    double* L, double* smax, double dt)
{
    const size_t numel = state_numel(p, p->meqn);
    int rc;
    double* qold = (double*)malloc(numel * sizeof(double));
    if (!qold) return FINESS_ERR_ARG;
    memcpy(qold, qnew, numel * sizeof(double));
    orc_set_bnd(p, qnew, aux);
    pif_flux_sum(p, qnew, aux, 1.0, 1, Fs, Gs);
    if ((rc = construct_l_noc(p, qnew, aux, L))) goto out;
    euler_step(p, 0.5 * dt, qold, L, qstar);
    orc_set_bnd(p, qstar, aux);
    pif_flux_sum(p, qstar, aux, 2.0, 0, Fs, Gs);
    if ((rc = construct_l_noc(p, qstar, aux, L))) goto out;
    euler_step(p, 0.5 * dt, qold, L, qstar);
    orc_set_bnd(p, qstar, aux);
    pif_flux_sum(p, qstar, aux, 2.0, 0, Fs, Gs);
    if ((rc = construct_l_noc(p, qstar, aux, L))) goto out;
    euler_step(p, dt, qold, L, qstar);
    orc_set_bnd(p, qstar, aux);
    pif_flux_sum(p, qstar, aux, 1.0, 0, Fs, Gs);
#pragma omp parallel for schedule(static)
    for (long long k = 0; k < (long long)numel; k++) {
        Fs[k] = Fs[k] / 6.0;
        Gs[k] = Gs[k] / 6.0;
    }
    orc_set_bnd(p, qnew, aux);
    if ((rc = construct_l_2d(p, qnew, aux, Fs, Gs, L, smax))) goto out;
    euler_step(p, dt, qold, L, qnew);
    orc_set_bnd(p, qnew, aux);
out:
    free(qold);
    return rc;
}

/* Shared controller of FinSolveRK / FinSolveLxW / FinSolveRK_PIF:
 * lib/2d/FinSolveRK.cpp:77-367, lib/2d/FinSolveLxW.cpp:100-229,
 * lib/2d/FinSolveRK_PIF.cpp:122-293, lib/1d/FinSolveRK.cpp:88-363. */
/* One step of the stepper `ts` (FINESS_TS_*). */
static int any_step(const finess_params* p, int ts, double* q, double* qstar, double* aux, double* F, double* G,
                    double* L, double* smax, double dt)
{
    if (ts == FINESS_TS_LXW) return lxw_step(p, q, aux, F, G, L, smax, dt);
    if (ts == FINESS_TS_PIF) return pif_step(p, q, qstar, aux, F, G, L, smax, dt);
    return rk_step(p, q, qstar, aux, L, smax, dt);
}

static int check_stepper(const finess_params* p, int ts)
{
    if (ts != FINESS_TS_RK && p->space_order != 5) return FINESS_ERR_CONFIG;
    if (ts == FINESS_TS_LXW && p->mbc < 5) return FINESS_ERR_CONFIG;
    if (ts == FINESS_TS_PIF && (p->ndims != 2 || p->time_order != 4)) return FINESS_ERR_CONFIG;
    /* source terms are restated for Runge-Kutta and ConstructL only (LxW: the
     * reference exits, lib/2d/ConstructLxWL.cpp:376-380) */
    if (ts != FINESS_TS_RK && p->source_term) return FINESS_ERR_CONFIG;
    return FINESS_OK;
}

static int fin_solve(const finess_params* p, int ts, double* qnew, double* aux, double* tio,
                     double tend, double* dtv, orc_stats* st)
{
    int rc = check_config(p);
    if (rc) return rc;
    if ((rc = check_stepper(p, ts))) return rc;
    const int lxw = ts != FINESS_TS_RK;  /* flux arrays needed */
    const size_t numel = state_numel(p, p->meqn);
    double* qold = (double*)malloc(numel * sizeof(double));
    double* qstar = (double*)malloc(numel * sizeof(double));
    double* L = (double*)calloc(numel, sizeof(double));
    double* smax = (double*)calloc(smax_numel(p), sizeof(double));
    double* F = lxw ? (double*)calloc(numel, sizeof(double)) : NULL;
    double* G = lxw ? (double*)calloc(numel, sizeof(double)) : NULL;
    const double CFL_max = p->max_cfl, CFL_target = p->desired_cfl;
    double t = *tio, dt = dtv[1], cfl = 0.0;
    double dtmin = dt, dtmax = dt;
    int n_step = 0;
    if (st) { memset(st, 0, sizeof(*st)); st->dt_min = dt; st->dt_max = dt; }
    memcpy(qold, qnew, numel * sizeof(double));
    memcpy(qstar, qnew, numel * sizeof(double));
    while (t < tend) {
        int m_accept = 0;
        n_step = n_step + 1;
        if (n_step > p->nv) { rc = FINESS_ERR_NV; goto done; }
        memcpy(qold, qnew, numel * sizeof(double));
        while (m_accept == 0) {
            const double told = t;
            if (told + dt > tend) dt = tend - told;
            t = told + dt;
            memset(smax, 0, smax_numel(p) * sizeof(double));
            rc = any_step(p, ts, qnew, qstar, aux, F, G, L, smax, dt);
            if (rc) goto done;
            rc = orc_get_cfl(p, dt, dtv[2], smax, &cfl);
            if (rc) goto done;
            if (cfl > 0.0) {
                dt = Min(dtv[2], dt * CFL_target / cfl);
                dtmin = Min(dt, dtmin);
                dtmax = Max(dt, dtmax);
            } else {
                dt = dtv[2];
            }
            if (cfl <= CFL_max) {
                m_accept = 1;
                if (st) st->steps_accepted++;
            } else {
                t = told;
                if (st) st->steps_rejected++;
                memcpy(qnew, qold, numel * sizeof(double));
            }
        }
        if (p->ndims == 1) orc_set_bnd(p, qnew, aux); /* lib/1d/FinSolveRK.cpp:353 */
    }
    dtv[1] = dt;
    rc = FINESS_OK;
done:
    *tio = t;
    if (st) { st->dt_min = dtmin; st->dt_max = dtmax; st->last_cfl = cfl; }
    free(qold); free(qstar); free(L); free(smax); free(F); free(G);
    return rc;
}

int orc_fin_solve_rk(const finess_params* p, double* q, double* aux, double* t, double tend,
                     double* dtv, orc_stats* st)
{
    return fin_solve(p, 0, q, aux, t, tend, dtv, st);
}

int orc_fin_solve_lxw(const finess_params* p, double* q, double* aux, double* t, double tend,
                      double* dtv, orc_stats* st)
{
    return fin_solve(p, FINESS_TS_LXW, q, aux, t, tend, dtv, st);
}

int orc_fin_solve_rk_pif(const finess_params* p, double* q, double* aux, double* t, double tend,
                         double* dtv, orc_stats* st)
{
    return fin_solve(p, FINESS_TS_PIF, q, aux, t, tend, dtv, st);
}

int orc_run_fixed_steps(const finess_params* p, double* q, double* aux, double dt, int nsteps, double* cfl)
{
    int rc = check_config(p);
    if (rc) return rc;
    const int ts = p->time_stepping_method;
    if ((rc = check_stepper(p, ts))) return rc;
    const int lxw = ts != FINESS_TS_RK;
    const size_t numel = state_numel(p, p->meqn);
    double* qstar = (double*)malloc(numel * sizeof(double));
    double* L = (double*)calloc(numel, sizeof(double));
    double* smax = (double*)calloc(smax_numel(p), sizeof(double));
    double* F = lxw ? (double*)calloc(numel, sizeof(double)) : NULL;
    double* G = lxw ? (double*)calloc(numel, sizeof(double)) : NULL;
    memcpy(qstar, q, numel * sizeof(double));
    double c = 0.0;
    for (int n = 0; n < nsteps && !rc; n++) {
        memset(smax, 0, smax_numel(p) * sizeof(double));
        rc = any_step(p, ts, q, qstar, aux, F, G, L, smax, dt);
        if (!rc) rc = orc_get_cfl(p, dt, 1e300, smax, &c);
    }
    if (cfl) *cfl = c;
    free(qstar); free(L); free(smax); free(F); free(G);
    return rc;
}

/* ====================================================================== */
/* Lax-Wendroff: time-integrated fluxes                                     */
/* ====================================================================== */

/* Diff1 / Diff2, lib/WenoReconstruct.cpp:702-716, 791-806 */
static double diff1(double dx, const double* f)
{
    double tmp = (f[0] - f[4]) * (1.0 / 12.0);
    tmp += (-f[1] + f[3]) * (2.0 / 3.0);
    return tmp / dx;
}
static double diff2(double dx, const double* f)
{
    double tmp = (-f[0] - f[4]) * (1.0 / 12.0);
    tmp += (f[1] + f[3]) * (4.0 / 3.0);
    tmp += (-f[2]) * (5.0 / 2.0);
    return tmp / (dx * dx);
}

/* ConstructIntegratedR (one-stage overload), lib/2d/ConstructIntegratedR.cpp:
 * 59-296: F = f + dt/2 (f_t + dt/3 f_tt) on cells [1-3, m+3]^2. */
/* 1D ConstructIntegratedF (lib/1d/ConstructIntegratedF.cpp:54-212):
 * F = f + dt/2 (f_t + dt/3 f_tt) on cells [1-3, mx+3] with f_t = -A f_x and,
 * for time_order > 2, f_tt = H(f_x, f_x) + A (H(q_x, f_x) + A f_xx); A and H
 * are the derivatives of FluxFunc (the advection flux aux(1) q: A = aux(1),
 * H = 0, apps/1d/advection/lib/{DFluxFunc,D2FluxFunc}.cpp). */
static int construct_integrated_f_1d(const finess_params* p, double dt, const double* q, const double* aux,
                                     double* F)
{
    const int mx = p->mx, meqn = p->meqn, mbc = p->mbc, na = naux(p);
    const double dx = p->dx;
    if (mbc < 5 || meqn != 1 || p->physics != FINESS_PHYS_ADVECTION) return FINESS_ERR_CONFIG;
    const int mbc_small = 3, half = 3, npts = 5;
    for (int i = 1 - mbc_small; i <= mx + mbc_small; i++) {
        double Fv[5], qv[5];
        int s = -half + 1;
        for (int r = 0; r < npts; r++, s++) {
            orc_flux(p, &q[ix2(i + s, 1, meqn, mbc)], &aux[ix2(i + s, 1, na, mbc)], &Fv[r], NULL);
            qv[r] = q[ix2(i + s, 1, meqn, mbc)];
        }
        const double fx = diff1(dx, Fv), qx = diff1(dx, qv), fxx = diff2(dx, Fv);
        (void)qx;  /* enters only through H, which is zero */
        const double A = aux[ix2(i, 1, na, mbc)];
        double ft = 0.;
        ft += -A * fx;
        double ftt = 0.;
        if (p->time_order > 2) {
            double tmp_vec = 0.;   /* sum of H q_x f_x terms: H = 0 */
            tmp_vec = tmp_vec + (0. + A * fxx);
            ftt = ftt + (0. + A * tmp_vec);
        }
        double f;
        orc_flux(p, &q[ix2(i, 1, meqn, mbc)], &aux[ix2(i, 1, na, mbc)], &f, NULL);
        F[ix2(i, 1, meqn, mbc)] = f + 0.5 * dt * (ft + dt / 3.0 * ftt);
    }
    return FINESS_OK;
}

int orc_construct_integrated_r(const finess_params* p, double dt, const double* q, const double* aux,
                               double* F, double* G)
{
    const int mx = p->mx, my = p->my, meqn = p->meqn, mbc = p->mbc, na = naux(p);
    const double dx = p->dx, dy = p->dy;
    if (p->ndims == 1) return construct_integrated_f_1d(p, dt, q, aux, F);
    if (mbc < 5 || p->ndims != 2) return FINESS_ERR_CONFIG;
    /* R = SampleFunction(FluxFunc) on all cells incl. ghosts, :83-84 */
    const size_t ncell = (size_t)(mx + 2 * mbc) * (my + 2 * mbc);
    double* R = (double*)malloc(ncell * meqn * 2 * sizeof(double));
    if (!R) return FINESS_ERR_ARG;
#define RIX(i, j, m, d) ((ix3((i), (j), 1, my, 1, mbc)) * (size_t)(meqn * 2) + (size_t)((m)-1) * 2 + (size_t)((d)-1))
#pragma omp parallel for schedule(static)
    for (int i = 1 - mbc; i <= mx + mbc; i++)
        for (int j = 1 - mbc; j <= my + mbc; j++) {
            double f[MAXEQ], g[MAXEQ];
            orc_flux(p, &q[ix3(i, j, 1, my, meqn, mbc)], &aux[ix3(i, j, 1, my, na, mbc)], f, g);
            for (int m = 1; m <= meqn; m++) {
                R[RIX(i, j, m, 1)] = f[m - 1];
                R[RIX(i, j, m, 2)] = g[m - 1];
            }
        }
    const int mbc_small = 3, half = 3, npts = 5;
#pragma omp parallel for schedule(static)
    for (int i = 1 - mbc_small; i <= mx + mbc_small; i++)
        for (int j = 1 - mbc_small; j <= my + mbc_small; j++) {
            double Fv[MAXEQ][5], Gv[MAXEQ][5], qx_[MAXEQ][5], qy_[MAXEQ][5];
            for (int m = 1; m <= meqn; m++) {
                int s = -half + 1;
                for (int r = 1; r <= npts; r++) {
                    Fv[m - 1][r - 1] = R[RIX(i + s, j, m, 1)];
                    Gv[m - 1][r - 1] = R[RIX(i, j + s, m, 2)];
                    qx_[m - 1][r - 1] = q[ix3(i + s, j, m, my, meqn, mbc)];
                    qy_[m - 1][r - 1] = q[ix3(i, j + s, m, my, meqn, mbc)];
                    s++;
                }
            }
            double fx[MAXEQ], qx[MAXEQ], gy[MAXEQ], qy[MAXEQ];
            for (int m = 0; m < meqn; m++) {
                fx[m] = diff1(dx, Fv[m]);
                qx[m] = diff1(dx, qx_[m]);
                gy[m] = diff1(dy, Gv[m]);
                qy[m] = diff1(dy, qy_[m]);
            }
            double A[2 * MAXEQ * MAXEQ];
            const double* qc = &q[ix3(i, j, 1, my, meqn, mbc)];
            orc_dflux(p, qc, A);
#define AA(m1, m2, d) A[(((d)-1) * meqn + (m1)-1) * meqn + (m2)-1]
            double f_t[MAXEQ], g_t[MAXEQ];
            for (int m1 = 1; m1 <= meqn; m1++) {
                double tmp1 = 0., tmp2 = 0.;
                for (int m2 = 1; m2 <= meqn; m2++) {
                    tmp1 += -(AA(m1, m2, 1)) * (fx[m2 - 1] + gy[m2 - 1]);
                    tmp2 += -(AA(m1, m2, 2)) * (fx[m2 - 1] + gy[m2 - 1]);
                }
                f_t[m1 - 1] = tmp1;
                g_t[m1 - 1] = tmp2;
            }
            double f_tt[MAXEQ] = {0}, g_tt[MAXEQ] = {0};
            if (p->time_order > 2) {
                double Hs[2 * MAXEQ * MAXEQ * MAXEQ];
                orc_d2flux(p, qc, Hs);
#define HH(m, m1, m2, d) Hs[((((d)-1) * meqn + (m)-1) * meqn + (m1)-1) * meqn + (m2)-1]
                double S[MAXEQ], St[MAXEQ] = {0};
                for (int m = 0; m < meqn; m++) S[m] = fx[m] + gy[m];
                double fxx[MAXEQ], gyy[MAXEQ], fxy[MAXEQ], gxy[MAXEQ];
                for (int m = 0; m < meqn; m++) {
                    fxx[m] = diff2(dx, Fv[m]);
                    gyy[m] = diff2(dy, Gv[m]);
                }
                for (int m = 1; m <= meqn; m++) { /* :202-213 */
                    double tmp1 = 0.5 * (R[RIX(i + 1, j + 1, m, 1)] - R[RIX(i - 1, j + 1, m, 1)]) / dx;
                    tmp1 -= 0.5 * (R[RIX(i + 1, j - 1, m, 1)] - R[RIX(i - 1, j - 1, m, 1)]) / dx;
                    tmp1 *= 0.5 / dy;
                    fxy[m - 1] = tmp1;
                    double tmp2 = 0.5 * (R[RIX(i + 1, j + 1, m, 2)] - R[RIX(i - 1, j + 1, m, 2)]) / dx;
                    tmp2 -= 0.5 * (R[RIX(i + 1, j - 1, m, 2)] - R[RIX(i - 1, j - 1, m, 2)]) / dx;
                    tmp2 *= 0.5 / dy;
                    gxy[m - 1] = tmp2;
                }
                for (int m = 1; m <= meqn; m++) { /* :215-240 */
                    double tmp = 0.;
                    for (int m1 = 1; m1 <= meqn; m1++)
                        for (int m2 = 1; m2 <= meqn; m2++) {
                            tmp += HH(m, m1, m2, 1) * qx[m1 - 1] * S[m2 - 1];
                            tmp += HH(m, m1, m2, 2) * qy[m1 - 1] * S[m2 - 1];
                        }
                    for (int m1 = 1; m1 <= meqn; m1++) {
                        tmp += AA(m, m1, 1) * (fxx[m1 - 1] + gxy[m1 - 1]);
                        tmp += AA(m, m1, 2) * (fxy[m1 - 1] + gyy[m1 - 1]);
                    }
                    St[m - 1] = tmp;
                }
                for (int m1 = 1; m1 <= meqn; m1++) { /* :243-256 */
                    double tmp1 = 0., tmp2 = 0.;
                    for (int m2 = 1; m2 <= meqn; m2++) {
                        tmp1 += AA(m1, m2, 1) * St[m2 - 1];
                        tmp2 += AA(m1, m2, 2) * St[m2 - 1];
                    }
                    f_tt[m1 - 1] = tmp1;
                    g_tt[m1 - 1] = tmp2;
                }
                for (int m = 1; m <= meqn; m++) { /* :259-272 */
                    double tmp1 = 0., tmp2 = 0.;
                    for (int m1 = 1; m1 <= meqn; m1++)
                        for (int m2 = 1; m2 <= meqn; m2++) {
                            tmp1 += HH(m, m1, m2, 1) * S[m1 - 1] * S[m2 - 1];
                            tmp2 += HH(m, m1, m2, 2) * S[m1 - 1] * S[m2 - 1];
                        }
                    f_tt[m - 1] = f_tt[m - 1] + tmp1;
                    g_tt[m - 1] = g_tt[m - 1] + tmp2;
                }
#undef HH
            }
#undef AA
            for (int m = 1; m <= meqn; m++) { /* :287-291 */
                F[ix3(i, j, m, my, meqn, mbc)] = R[RIX(i, j, m, 1)] + 0.5 * dt * (f_t[m - 1] + dt / 3.0 * f_tt[m - 1]);
                G[ix3(i, j, m, my, meqn, mbc)] = R[RIX(i, j, m, 2)] + 0.5 * dt * (g_t[m - 1] + dt / 3.0 * g_tt[m - 1]);
            }
        }
#undef RIX
    free(R);
    return FINESS_OK;
}

/* ====================================================================== */
/* Troubled-interface indicator (new functionality, no reference code)      */
/* ====================================================================== */

/* Rotated primitives (rho, u_n, u_t, p) of an Euler state. */
static void euler_prims(const finess_params* p, int ixy, const double* q, double* v)
{
    const double rho = q[0];
    const double un = (ixy == 1 ? q[1] : q[2]) / rho, ut = (ixy == 1 ? q[2] : q[1]) / rho, uw = q[3] / rho;
    v[0] = rho;
    v[1] = un;
    v[2] = ut;
    v[3] = (p->gamma - 1.0) * (q[4] - 0.5 * rho * (un * un + ut * ut + uw * uw));
}

void orc_face_features(const finess_params* p, int ixy, const double* qs, float* feat)
{
    double v[6][4], mean[4], lo[4], hi[4];
    for (int s = 0; s < 6; s++) euler_prims(p, ixy, qs + s * p->meqn, v[s]);
    for (int k = 0; k < 4; k++) {
        mean[k] = 0.0;
        lo[k] = hi[k] = v[0][k];
        for (int s = 0; s < 6; s++) {
            mean[k] += v[s][k];
            if (v[s][k] < lo[k]) lo[k] = v[s][k];
            if (v[s][k] > hi[k]) hi[k] = v[s][k];
        }
        mean[k] /= 6.0;
    }
    const double cref = sqrt(fabs(p->gamma * mean[3] / mean[0]));
    const double ref[4] = {fabs(mean[0]), cref, cref, fabs(mean[3])};
    for (int k = 0; k < 4; k++)
        for (int s = 0; s < 6; s++) feat[s * 4 + k] = (float)((v[s][k] - mean[k]) / (hi[k] - lo[k] + 0.01 * ref[k] + 1e-300));
}

/* The reference's split characteristic stencils (lib/2d/ConstructL.cpp:
 * 72-185) and the JS weights of each (lib/WenoReconstruct.cpp:131-134). */
double orc_face_departure(const finess_params* p, int ixy, const double* qs)
{
    const int meqn = p->meqn;
    double Qavg[MAXEQ], qv[MAXEQ * 6], fv[MAXEQ * 6], wv[MAXEQ * 6], gv[MAXEQ * 6], F[MAXEQ], G[MAXEQ], aux[1] = {0};
    for (int m = 0; m < meqn; m++) Qavg[m] = 0.5 * (qs[3 * meqn + m] + qs[2 * meqn + m]);
    for (int s = 0; s < 6; s++) {
        orc_flux(p, qs + s * meqn, aux, F, G);
        for (int m = 0; m < meqn; m++) {
            qv[m * 6 + s] = qs[s * meqn + m];
            fv[m * 6 + s] = ixy == 1 ? F[m] : G[m];
        }
    }
    orc_project_left(p, ixy, Qavg, 6, qv, wv);
    orc_project_left(p, ixy, Qavg, 6, fv, gv);
    double s1, s2;
    orc_wave_speed(p, ixy, qs + 2 * meqn, qs + 3 * meqn, aux, aux, &s1, &s2);
    const double la = p->alpha_scaling * Max(fabs(s1), fabs(s2));
    const double d[3] = {0.1, 0.6, 0.3};
    double dep = 0.0;
    for (int m = 0; m < meqn; m++) {
        double gp[5], gm[5], w[3];
        for (int s = 1; s <= 5; s++) {
            gp[s - 1] = 0.5 * (gv[m * 6 + s - 1] + la * wv[m * 6 + s - 1]);
            gm[s - 1] = 0.5 * (gv[m * 6 + 6 - s] - la * wv[m * 6 + 6 - s]);
        }
        orc_weno5_js_weights(p->epsilon, 2.0, gp, w);
        for (int j = 0; j < 3; j++) dep = Max(dep, fabs(w[j] - d[j]));
        orc_weno5_js_weights(p->epsilon, 2.0, gm, w);
        for (int j = 0; j < 3; j++) dep = Max(dep, fabs(w[j] - d[j]));
    }
    return dep;
}

/* Component-wise LLF with the linear fifth-order reconstruction
 * (WenoReconstruct_FD5 on the conserved components, identity projection). */
static double interface_flux_linear(const finess_params* p, int ixy, const double qs[6][MAXEQ],
                                    const double fs[6][MAXEQ], double* fhat)
{
    double s1, s2, aux[1] = {0};
    orc_wave_speed(p, ixy, qs[2], qs[3], aux, aux, &s1, &s2);
    const double alpha = Max(0.0, Max(fabs(s1), fabs(s2)));
    const double la = p->alpha_scaling * alpha;
    for (int m = 0; m < p->meqn; m++) {
        double gp[5], gm[5];
        for (int s = 1; s <= 5; s++) {
            gp[s - 1] = 0.5 * (fs[s - 1][m] + la * qs[s - 1][m]);
            gm[s - 1] = 0.5 * (fs[6 - s][m] - la * qs[6 - s][m]);
        }
        fhat[m] = orc_weno5(FINESS_WENO_FD, p->epsilon, p->power_param, gp) +
                  orc_weno5(FINESS_WENO_FD, p->epsilon, p->power_param, gm);
    }
    return alpha;
}

double orc_face_flux_linear(const finess_params* p, int ixy, const double* qs, double* fhat)
{
    double q6[6][MAXEQ], f6[6][MAXEQ], F[MAXEQ], G[MAXEQ], aux[1] = {0};
    for (int s = 0; s < 6; s++) {
        for (int m = 0; m < p->meqn; m++) q6[s][m] = qs[s * p->meqn + m];
        orc_flux(p, q6[s], aux, F, G);
        for (int m = 0; m < p->meqn; m++) f6[s][m] = ixy == 1 ? F[m] : G[m];
    }
    return interface_flux_linear(p, ixy, q6, f6, fhat);
}

/* mask (may be NULL): per-face routing, x faces (i-1)*my + j-1 then y faces
 * (mx+1)*my + (i-1)*(my+1) + j-1; mask_read: take mode-2 decisions from it,
 * else record them there. */
static int construct_l_routed_mask(const finess_params* p, const double* q, const double* aux, int mode,
                                   double tau, unsigned char* mask, int mask_read, double* Lstar, double* smax,
                                   long long* n_troubled)
{
    const int mx = p->mx, my = p->my, meqn = p->meqn, mbc = p->mbc, na = naux(p);
    if (p->ndims != 2 || p->physics != FINESS_PHYS_EULER || check_config5(p) || p->global_alpha)
        return FINESS_ERR_CONFIG;
    double* Fhat = (double*)calloc((size_t)(mx + 1 + 2 * mbc) * (my + 2 * mbc) * meqn, sizeof(double));
    double* Ghat = (double*)calloc((size_t)(mx + 2 * mbc) * (my + 1 + 2 * mbc) * meqn, sizeof(double));
    long long nt = 0;
    for (int dir = 1; dir <= 2; dir++) {
        const int imax = dir == 1 ? mx + 1 : mx, jmax = dir == 1 ? my : my + 1;
#pragma omp parallel for schedule(static) reduction(+ : nt)
        for (int i = 1; i <= imax; i++)
            for (int j = 1; j <= jmax; j++) {
                double qs[6][MAXEQ], fs[6][MAXEQ], F[MAXEQ], G[MAXEQ], fh[MAXEQ];
                for (int s = 1; s <= 6; s++) {
                    const int is = dir == 1 ? i - 1 + s - 3 : i, js = dir == 1 ? j : j - 1 + s - 3;
                    for (int m = 1; m <= meqn; m++) qs[s - 1][m - 1] = q[ix3(is, js, m, my, meqn, mbc)];
                    orc_flux(p, qs[s - 1], &aux[ix3(is, js, 1, my, na, mbc)], F, G);
                    for (int m = 0; m < meqn; m++) fs[s - 1][m] = dir == 1 ? F[m] : G[m];
                }
                int full = mode == 0;
                const long long mi = dir == 1 ? (long long)(i - 1) * my + (j - 1)
                                              : (long long)(mx + 1) * my + (long long)(i - 1) * (my + 1) + (j - 1);
                if (mode == 2) {
                    if (mask && mask_read) full = mask[mi] != 0;
                    else full = orc_face_departure(p, dir, &qs[0][0]) > tau;
                    if (mask && !mask_read) mask[mi] = (unsigned char)full;
                }
                nt += full;
                double alpha;
                if (full) {
                    const int il = dir == 1 ? i - 1 : i, jl = dir == 1 ? j : j - 1;
                    alpha = interface_flux(p, dir, qs, fs, &aux[ix3(il, jl, 1, my, na, mbc)],
                                           &aux[ix3(i, j, 1, my, na, mbc)], 0.0, fh);
                } else {
                    alpha = interface_flux_linear(p, dir, qs, fs, fh);
                }
                double* sm = &smax[ix3(i, j, dir, my + 1, 2, mbc)];
                *sm = Max(*sm, alpha);
                for (int m = 1; m <= meqn; m++) {
                    if (dir == 1) Fhat[ix3(i, j, m, my, meqn, mbc)] = fh[m - 1];
                    else Ghat[ix3(i, j, m, my + 1, meqn, mbc)] = fh[m - 1];
                }
            }
    }
    for (int i = 1; i <= mx; i++)
        for (int j = 1; j <= my; j++)
            for (int m = 1; m <= meqn; m++) {
                double tmp = -(Fhat[ix3(i + 1, j, m, my, meqn, mbc)] - Fhat[ix3(i, j, m, my, meqn, mbc)]) / p->dx;
                tmp = tmp - (Ghat[ix3(i, j + 1, m, my + 1, meqn, mbc)] - Ghat[ix3(i, j, m, my + 1, meqn, mbc)]) / p->dy;
                Lstar[ix3(i, j, m, my, meqn, mbc)] = tmp;
            }
    free(Fhat);
    free(Ghat);
    if (n_troubled) *n_troubled = nt;
    return FINESS_OK;
}

int orc_construct_l_routed(const finess_params* p, const double* q, const double* aux, int mode, double tau,
                           double* Lstar, double* smax, long long* n_troubled)
{
    return construct_l_routed_mask(p, q, aux, mode, tau, NULL, 0, Lstar, smax, n_troubled);
}

int orc_run_fixed_steps_routed(const finess_params* p, double* q, double* aux, double dt, int nsteps, double tau,
                               int per_step, double* cfl)
{
    int rc = check_config5(p);
    if (rc) return rc;
    if (p->ndims != 2 || p->time_stepping_method != FINESS_TS_RK || p->time_order != 3) return FINESS_ERR_CONFIG;
    const size_t numel = state_numel(p, p->meqn);
    double* qstar = (double*)malloc(numel * sizeof(double));
    double* L = (double*)calloc(numel, sizeof(double));
    double* smax = (double*)calloc(smax_numel(p), sizeof(double));
    unsigned char* mask = (unsigned char*)calloc((size_t)(p->mx + 1) * p->my + (size_t)p->mx * (p->my + 1), 1);
    const int k = 2;
    double c = 0.0;
    for (int n = 0; n < nsteps && !rc; n++) {
        memset(smax, 0, smax_numel(p) * sizeof(double));
        unsigned char* m = per_step ? mask : NULL;
        orc_set_bnd(p, q, aux);
        /* the first stage decides at every per_step-th step, else reuses */
        const int reuse = per_step >= 1 && n % per_step != 0;
        rc = construct_l_routed_mask(p, q, aux, 2, tau, m, reuse, L, smax, NULL);
        if (rc) break;
        update_soln(numel, RK_A1[k][0], RK_A2[k][0], RK_B[k][0], dt, q, L, qstar);
        orc_set_bnd(p, qstar, aux);
        rc = construct_l_routed_mask(p, qstar, aux, 2, tau, m, 1, L, smax, NULL);
        if (rc) break;
        update_soln(numel, RK_A1[k][1], RK_A2[k][1], RK_B[k][1], dt, q, L, qstar);
        orc_set_bnd(p, qstar, aux);
        rc = construct_l_routed_mask(p, qstar, aux, 2, tau, m, 1, L, smax, NULL);
        if (rc) break;
        update_soln(numel, RK_A1[k][2], RK_A2[k][2], RK_B[k][2], dt, qstar, L, q);
        rc = orc_get_cfl(p, dt, 1e300, smax, &c);
    }
    if (cfl) *cfl = c;
    free(qstar); free(L); free(smax); free(mask);
    return rc;
}

/* ---- MLP with BF16 operand rounding ------------------------------------- */
#define NIN 24
#define NH 32
#define OFF_W1 0
#define OFF_B1 (OFF_W1 + NIN * NH)
#define OFF_W2 (OFF_B1 + NH)
#define OFF_B2 (OFF_W2 + NH * NH)
#define OFF_W3 (OFF_B2 + NH)
#define OFF_B3 (OFF_W3 + NH)
#define NPARAM (OFF_B3 + 1)

float orc_bf16(float x)
{
    uint32_t u;
    memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    float r;
    memcpy(&r, &u, 4);
    return r;
}
static int g_round_bf16 = 1;
void orc_mlp_set_rounding(int on) { g_round_bf16 = on; }
static double rb(double x) { return g_round_bf16 ? (double)orc_bf16((float)x) : x; }

/* One sample: activations in FP32 as the GPU stores them. */
static void mlp_sample(const float* P, const float* x, float* h1, float* h2, float* z)
{
    for (int n = 0; n < NH; n++) {
        double a = 0.0;
        for (int k = 0; k < NIN; k++) a += rb(x[k]) * rb(P[OFF_W1 + k * NH + n]);
        const float v = (float)a + P[OFF_B1 + n];
        h1[n] = v > 0.f ? v : 0.f;
    }
    for (int n = 0; n < NH; n++) {
        double a = 0.0;
        for (int k = 0; k < NH; k++) a += rb(h1[k]) * rb(P[OFF_W2 + k * NH + n]);
        const float v = (float)a + P[OFF_B2 + n];
        h2[n] = v > 0.f ? v : 0.f;
    }
    double a = 0.0;
    for (int k = 0; k < NH; k++) a += rb(h2[k]) * rb(P[OFF_W3 + k]);
    *z = (float)a + P[OFF_B3];
}

void orc_mlp_forward(const float* P, const float* X, int n, float* logits)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        float h1[NH], h2[NH];
        mlp_sample(P, X + (size_t)i * NIN, h1, h2, &logits[i]);
    }
}

double orc_mlp_gradient(const float* P, const float* X, const float* y, int n, double* grad)
{
    memset(grad, 0, NPARAM * sizeof(double));
    double loss = 0.0;
    for (int i = 0; i < n; i++) {
        const float* x = X + (size_t)i * NIN;
        float h1[NH], h2[NH], z;
        mlp_sample(P, x, h1, h2, &z);
        const double s = 1.0 / (1.0 + exp(-(double)z));
        loss += fmax((double)z, 0.0) - z * (double)y[i] + log1p(exp(-fabs((double)z)));
        const float dz = (float)((s - y[i]) / n);
        /* gW3 = H2^T dz, gb3 = sum dz */
        for (int k = 0; k < NH; k++) grad[OFF_W3 + k] += rb(h2[k]) * rb(dz);
        grad[OFF_B3] += dz;
        /* dH2 = dz w3^T . [H2 > 0] */
        float dh2[NH], dh1[NH];
        for (int k = 0; k < NH; k++) dh2[k] = h2[k] > 0.f ? (float)(rb(dz) * rb(P[OFF_W3 + k])) : 0.f;
        for (int k = 0; k < NH; k++)
            for (int m = 0; m < NH; m++) grad[OFF_W2 + k * NH + m] += rb(h1[k]) * rb(dh2[m]);
        for (int m = 0; m < NH; m++) grad[OFF_B2 + m] += dh2[m];
        /* dH1 = dH2 W2^T . [H1 > 0] */
        for (int k = 0; k < NH; k++) {
            double a = 0.0;
            for (int m = 0; m < NH; m++) a += rb(dh2[m]) * rb(P[OFF_W2 + k * NH + m]);
            dh1[k] = h1[k] > 0.f ? (float)a : 0.f;
        }
        for (int k = 0; k < NIN; k++)
            for (int m = 0; m < NH; m++) grad[OFF_W1 + k * NH + m] += rb(x[k]) * rb(dh1[m]);
        for (int m = 0; m < NH; m++) grad[OFF_B1 + m] += dh1[m];
    }
    return loss / n;
}

int orc_construct_lxw_l(const finess_params* p, const double* q, const double* aux, const double* F,
                        const double* G, double* Lstar, double* smax)
{
    const int rc = check_config5(p);
    if (rc) return rc;
    return construct_l_2d(p, q, aux, F, G, Lstar, smax);
}
