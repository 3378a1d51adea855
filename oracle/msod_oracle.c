/*
 * CPU oracle: a plain-C restatement of the miniSOD2D hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Written from the reference's Fortran (cited
 * file:line per function, paths relative to the reference's src/); it follows
 * the reference's loop structure and operation order so that it agrees with
 * the reference to rounding level.  It is the checker for the HIP kernels and
 * is never linked into the product.
 */
#include "msod_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* GLL nodes and weights                                                     */
/* ------------------------------------------------------------------------ */

/* Legendre polynomial of degree n and its first two derivatives at x, by the
 * three-term recurrence (utils/sources/mod_maths.f90:113-152, VALEPO). */
static void valepo(int n, double x, double *y, double *dy, double *d2y)
{
    *y = 1.0; *dy = 0.0; *d2y = 0.0;
    if (n == 0) return;
    *y = x; *dy = 1.0; *d2y = 0.0;
    if (n == 1) return;
    double yp = 1.0, dyp = 0.0, d2yp = 0.0;
    for (int i = 2; i <= n; ++i) {
        double c1 = (double)i, c2 = 2.0 * c1 - 1.0, c4 = c1 - 1.0;
        double ym = *y;
        *y = (c2 * x * (*y) - c4 * yp) / c1;
        yp = ym;
        double dym = *dy;
        *dy = (c2 * x * (*dy) - c4 * dyp + c2 * yp) / c1;
        dyp = dym;
        double d2ym = *d2y;
        *d2y = (c2 * x * (*d2y) - c4 * d2yp + 2.0 * c2 * dyp) / c1;
        d2yp = d2ym;
    }
}

/* Roots of (1-x^2)P'_n, ascending, with Newton (8 steps) from Chebyshev
 * guesses (mod_maths.f90:46-82, ZELEGL); vn = P_n at the nodes. */
static void zelegl(int n, double *et, double *vn)
{
    if (n == 0) return;
    int n2 = (n - 1) / 2;
    double sn = (double)(2 * n - 4 * n2 - 3);
    et[0] = -1.0; et[n] = 1.0; vn[0] = sn; vn[n] = 1.0;
    if (n == 1) return;
    double y, dy, d2y;
    et[n2 + 1] = 0.0;
    valepo(n, 0.0, &y, &dy, &d2y);
    vn[n2 + 1] = y;
    if (n == 2) return;
    const double v_pi = 2.0 * asin(1.0);   /* mod_constants.f90:24 at rp=8 */
    double c = v_pi / (double)n;
    for (int i = 1; i <= n2; ++i) {
        double etx = cos(c * (double)i);
        for (int it = 0; it < 8; ++it) {
            valepo(n, etx, &y, &dy, &d2y);
            etx = etx - dy / d2y;
        }
        et[i] = -etx; et[n - i] = etx;
        vn[i] = y * sn; vn[n - i] = y;
    }
}

/* Weights 2/(n(n+1) P_n^2) (mod_maths.f90:84-111, WELEGL). */
static void welegl(int n, const double *et, const double *vn, double *wt)
{
    (void)et;
    if (n == 0) return;
    int n2 = (n - 1) / 2;
    double dn = (double)n;
    double c = 2.0 / (dn * (dn + 1.0));
    for (int i = 0; i <= n2; ++i) {
        double yv = vn[i];
        double w = c / (yv * yv);
        wt[i] = w; wt[n - i] = w;
    }
    if (n - 1 == 2 * n2) return;
    double yv = vn[n2 + 1];
    wt[n2 + 1] = c / (yv * yv);
}

void msod_oracle_gll_double(int p, double *xi, double *w)
{
    double vn[32];
    zelegl(p, xi, vn);
    welegl(p, xi, vn, w);
}

/* The reference rounds roots and weights through default REAL (binary32)
 * with a kind-less real() (mod_maths.f90:25,39-40), then widens them. */
void msod_oracle_gll(int p, double *xi, double *w)
{
    msod_oracle_gll_double(p, xi, w);
    for (int i = 0; i <= p; ++i) {
        xi[i] = (double)(float)xi[i];
        w[i] = (double)(float)w[i];
    }
}

/* ------------------------------------------------------------------------ */
/* Lagrange basis in SOD node order                                          */
/* ------------------------------------------------------------------------ */

/* lorder: SOD 1-D index -> position in the ascending grid, 0-based here:
 * [0, p, 1, 2, ..., p-1] (mod_maths.f90:199-203). */
static void make_lorder(int p, int *lo)
{
    lo[0] = 0;
    if (p >= 1) lo[1] = p;
    for (int i = 2; i <= p; ++i) lo[i] = i - 1;
}

/* mod_maths.f90:177-213 */
static void lagrange_poly(int p, const double *xi, double x, double *l)
{
    int lo[32];
    make_lorder(p, lo);
    for (int i = 0; i <= p; ++i) {
        l[i] = 1.0;
        for (int j = 0; j <= p; ++j)
            if (j != lo[i]) l[i] = l[i] * ((x - xi[j]) / (xi[lo[i]] - xi[j]));
    }
}

/* mod_maths.f90:215-257 */
static void lagrange_deriv(int p, const double *xi, double x, double *dl)
{
    int lo[32];
    make_lorder(p, lo);
    for (int i = 0; i <= p; ++i) {
        dl[i] = 0.0;
        for (int j = 0; j <= p; ++j) {
            double aux = 1.0;
            if (j != lo[i]) {
                for (int m = 0; m <= p; ++m)
                    if (m != j && m != lo[i]) aux = aux * (x - xi[m]) / (xi[lo[i]] - xi[m]);
                dl[i] = dl[i] + (1.0 / (xi[lo[i]] - xi[j])) * aux;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Gmsh high-order node tables                                               */
/* ------------------------------------------------------------------------ */

/* fem/sources/elem_qua.f90:305 and elem_hex.f90:119-120 (1-based corners) */
static const int QUAD_EDGES[4][2] = {{1, 2}, {2, 3}, {3, 4}, {4, 1}};
static const int HEX_EDGES[12][2] = {{1, 2}, {1, 4}, {1, 5}, {2, 3}, {2, 6}, {3, 4},
                                     {3, 7}, {4, 8}, {5, 6}, {5, 8}, {6, 7}, {7, 8}};
static const int HEX_FACES[6][4] = {{1, 4, 3, 2}, {1, 2, 6, 5}, {1, 5, 8, 4},
                                    {2, 3, 7, 6}, {3, 4, 8, 7}, {5, 6, 7, 8}};

/* fem/sources/mod_gmsh_indices.f90:66-94; tab[(p+1)^2][2] row-major here */
static void gen_quad(int p, int *tab)
{
    tab[0] = 0; tab[1] = 0;
    if (p <= 0) return;
    int c[4][2] = {{0, 0}, {p, 0}, {p, p}, {0, p}};
    for (int r = 1; r < 4; ++r) { tab[2 * r] = c[r][0]; tab[2 * r + 1] = c[r][1]; }
    if (p <= 1) return;
    int inode = 4;
    for (int ie = 0; ie < 4; ++ie) {
        int i0 = QUAD_EDGES[ie][0] - 1, i1 = QUAD_EDGES[ie][1] - 1;
        int u0 = (tab[2 * i1] - tab[2 * i0]) / p, u1 = (tab[2 * i1 + 1] - tab[2 * i0 + 1]) / p;
        for (int i = 1; i <= p - 1; ++i) {
            tab[2 * inode] = tab[2 * i0] + i * u0;
            tab[2 * inode + 1] = tab[2 * i0 + 1] + i * u1;
            ++inode;
        }
    }
    int nf = (p - 1) * (p - 1);
    int *tf = (int *)malloc(sizeof(int) * 2 * (size_t)nf);
    gen_quad(p - 2, tf);
    for (int i = 0; i < nf; ++i) {
        tab[2 * inode] = tf[2 * i] + 1;
        tab[2 * inode + 1] = tf[2 * i + 1] + 1;
        ++inode;
    }
    free(tf);
}

/* fem/sources/mod_gmsh_indices.f90:11-64; tab[(p+1)^3][3] row-major here */
static void gen_hex(int p, int *tab)
{
    tab[0] = tab[1] = tab[2] = 0;
    if (p <= 0) return;
    int c[8][3] = {{0, 0, 0}, {p, 0, 0}, {p, p, 0}, {0, p, 0},
                   {0, 0, p}, {p, 0, p}, {p, p, p}, {0, p, p}};
    for (int r = 1; r < 8; ++r)
        for (int d = 0; d < 3; ++d) tab[3 * r + d] = c[r][d];
    if (p <= 1) return;
    int inode = 8;
    for (int ie = 0; ie < 12; ++ie) {
        int i0 = HEX_EDGES[ie][0] - 1, i1 = HEX_EDGES[ie][1] - 1;
        int u[3];
        for (int d = 0; d < 3; ++d) u[d] = (tab[3 * i1 + d] - tab[3 * i0 + d]) / p;
        for (int i = 1; i <= p - 1; ++i) {
            for (int d = 0; d < 3; ++d) tab[3 * inode + d] = tab[3 * i0 + d] + i * u[d];
            ++inode;
        }
    }
    int nf = (p - 1) * (p - 1);
    int *tf = (int *)malloc(sizeof(int) * 2 * (size_t)nf);
    gen_quad(p - 2, tf);
    for (int i = 0; i < 2 * nf; ++i) tf[i] += 1;
    for (int f = 0; f < 6; ++f) {
        int i0 = HEX_FACES[f][0] - 1, i1 = HEX_FACES[f][1] - 1, i3 = HEX_FACES[f][3] - 1;
        int u[3], v[3];
        for (int d = 0; d < 3; ++d) {
            u[d] = (tab[3 * i1 + d] - tab[3 * i0 + d]) / p;
            v[d] = (tab[3 * i3 + d] - tab[3 * i0 + d]) / p;
        }
        for (int i = 0; i < nf; ++i) {
            for (int d = 0; d < 3; ++d)
                tab[3 * inode + d] = tab[3 * i0 + d] + u[d] * tf[2 * i] + v[d] * tf[2 * i + 1];
            ++inode;
        }
    }
    free(tf);
    int nv = (p - 1) * (p - 1) * (p - 1);
    int *tv = (int *)malloc(sizeof(int) * 3 * (size_t)nv);
    gen_hex(p - 2, tv);
    for (int i = 0; i < nv; ++i) {
        for (int d = 0; d < 3; ++d) tab[3 * inode + d] = tv[3 * i + d] + 1;
        ++inode;
    }
    free(tv);
}

/* Driver tables: app/sources/main.f90:110-255, quadrature_rules.f90:50-82,
 * elem_hex.f90:124-146, mod_maths.f90:259-297. */
void msod_oracle_tables(int p, int32_t *gmshIJK, int32_t *invAtoIJK, int32_t *AtoIJK,
                        int32_t *AtoI, int32_t *AtoJ, int32_t *AtoK, double *xgp, double *wgp,
                        double *Ngp, double *dNgp, double *dlxigp_ip)
{
    const int q = p + 1, nn = q * q * q;
    int *tab = (int *)malloc(sizeof(int) * 3 * (size_t)nn);
    gen_hex(p, tab);
    /* gmsh 1-D position -> SOD 1-D index (main.f90:133-157) */
    for (int a = 0; a < nn; ++a)
        for (int d = 0; d < 3; ++d) {
            int v = tab[3 * a + d];
            tab[3 * a + d] = (v == 0) ? 1 : (v == p ? 2 : v + 2);
        }
    for (int a = 0; a < nn; ++a)
        for (int d = 0; d < 3; ++d) gmshIJK[a + nn * d] = tab[3 * a + d];
    /* invAtoIJK and AtoIJK in the driver's (k,i,j) order (main.f90:168-185) */
    int jn = 0;
    for (int k = 1; k <= q; ++k)
        for (int i = 1; i <= q; ++i)
            for (int j = 1; j <= q; ++j) {
                int found = 0;
                for (int a = 0; a < nn; ++a)
                    if (tab[3 * a] == i && tab[3 * a + 1] == j && tab[3 * a + 2] == k) { found = a + 1; break; }
                invAtoIJK[(i - 1) + q * ((j - 1) + q * (k - 1))] = found;
                AtoIJK[jn++] = found;
            }
    for (int a = 0; a < nn; ++a) { AtoI[a] = tab[3 * a]; AtoJ[a] = tab[3 * a + 1]; AtoK[a] = tab[3 * a + 2]; }
    free(tab);

    /* GLL hex rule (quadrature_rules.f90:50-82) */
    double xi[32], w1[32];
    int lo[32];
    msod_oracle_gll(p, xi, w1);
    make_lorder(p, lo);
    int in = 0;
    for (int k = 0; k < q; ++k)
        for (int i = 0; i < q; ++i)
            for (int j = 0; j < q; ++j) {
                int a = AtoIJK[in++] - 1;
                xgp[a] = xi[lo[i]];
                xgp[a + nn] = xi[lo[j]];
                xgp[a + 2 * nn] = xi[lo[k]];
                wgp[a] = w1[lo[i]] * w1[lo[j]] * w1[lo[k]];
            }

    /* shape functions at every Gauss point (elem_hex.f90:124-146 with the
     * GLL grid; TripleTensorProduct, mod_maths.f90:259-297) */
    double l1[32], l2[32], l3[32], d1[32], d2[32], d3[32];
    for (int g = 0; g < nn; ++g) {
        double s = xgp[g], t = xgp[g + nn], z = xgp[g + 2 * nn];
        lagrange_poly(p, xi, s, l1);
        lagrange_poly(p, xi, t, l2);
        lagrange_poly(p, xi, z, l3);
        lagrange_deriv(p, xi, s, d1);
        lagrange_deriv(p, xi, t, d2);
        lagrange_deriv(p, xi, z, d3);
        for (int n = 0; n < q; ++n) {
            dlxigp_ip[g + nn * (0 + 3 * n)] = d1[n];
            dlxigp_ip[g + nn * (1 + 3 * n)] = d2[n];
            dlxigp_ip[g + nn * (2 + 3 * n)] = d3[n];
        }
        int c = 0;
        for (int k = 0; k < q; ++k)
            for (int i = 0; i < q; ++i)
                for (int j = 0; j < q; ++j) {
                    int a = AtoIJK[c++] - 1;
                    Ngp[g + nn * a] = l1[i] * l2[j] * l3[k];
                    dNgp[0 + 3 * (a + nn * g)] = d1[i] * l2[j] * l3[k];
                    dNgp[1 + 3 * (a + nn * g)] = l1[i] * d2[j] * l3[k];
                    dNgp[2 + 3 * (a + nn * g)] = l1[i] * l2[j] * d3[k];
                }
    }
}

/* ------------------------------------------------------------------------ */
/* Jacobian: fem/sources/jacobian_oper.f90:7-136                             */
/* ------------------------------------------------------------------------ */
void msod_oracle_elem_jacobian(int p, int nelem, int npoin, const int32_t *connec,
                               const double *coord, const double *dNgp, const double *wgp,
                               double *gpvol, double *He)
{
    const int q = p + 1, nn = q * q * q;
    for (long e = 0; e < nelem; ++e) {
        for (int g = 0; g < nn; ++g) {
            double Je[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    double s = 0.0;
                    for (int a = 0; a < nn; ++a) {
                        long ip = connec[e + (long)nelem * a] - 1;
                        s += dNgp[i + 3 * (a + nn * g)] * coord[ip + (long)npoin * j];
                    }
                    Je[i][j] = s;
                }
            double det = Je[0][0] * Je[1][1] * Je[2][2] + Je[0][1] * Je[1][2] * Je[2][0] +
                         Je[0][2] * Je[1][0] * Je[2][1] - Je[2][0] * Je[1][1] * Je[0][2] -
                         Je[2][1] * Je[1][2] * Je[0][0] - Je[2][2] * Je[1][0] * Je[0][1];
            double a[9], b[9];
            a[0] = Je[1][1] * Je[2][2] - Je[2][1] * Je[1][2];
            a[1] = Je[1][0] * Je[2][2] - Je[2][0] * Je[1][2];
            a[2] = Je[1][0] * Je[2][1] - Je[2][0] * Je[1][1];
            a[3] = Je[0][1] * Je[2][2] - Je[2][1] * Je[0][2];
            a[4] = Je[0][0] * Je[2][2] - Je[2][0] * Je[0][2];
            a[5] = Je[0][0] * Je[2][1] - Je[2][0] * Je[0][1];
            a[6] = Je[0][1] * Je[1][2] - Je[1][1] * Je[0][2];
            a[7] = Je[0][0] * Je[1][2] - Je[1][0] * Je[0][2];
            a[8] = Je[0][0] * Je[1][1] - Je[1][0] * Je[0][1];
            a[1] = -a[1]; a[3] = -a[3]; a[5] = -a[5]; a[7] = -a[7];
            for (int i = 0; i < 9; ++i) b[i] = a[i];
            b[1] = a[3]; b[2] = a[6]; b[3] = a[1]; b[5] = a[7]; b[6] = a[2]; b[7] = a[5];
            for (int i = 0; i < 9; ++i) b[i] = (1.0 / det) * b[i];
            double *H = He + 9 * ((long)g + (long)nn * e);
            /* He(1,1)=b1 He(1,2)=b2 He(1,3)=b3 He(2,1)=b4 ... (column-major) */
            H[0 + 3 * 0] = b[0]; H[0 + 3 * 1] = b[1]; H[0 + 3 * 2] = b[2];
            H[1 + 3 * 0] = b[3]; H[1 + 3 * 1] = b[4]; H[1 + 3 * 2] = b[5];
            H[2 + 3 * 0] = b[6]; H[2 + 3 * 1] = b[7]; H[2 + 3 * 2] = b[8];
            gpvol[g + (long)nn * e] = wgp[g] * det;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Convection                                                                */
/* ------------------------------------------------------------------------ */

#ifdef _OPENMP
#define ATOMIC_ADD(dst, v) do { _Pragma("omp atomic") (dst) += (v); } while (0)
#else
#define ATOMIC_ADD(dst, v) do { (dst) += (v); } while (0)
#endif

/* Rmass, Rmom: kernels/sources/elem_convec.f90:228-428 (fused kernel).
 * Rener: elem_convec.f90:676-831 (energy kernel of the split form, where
 * fuel = E*u and fukl = k*u are defined, :694-697). */
void msod_oracle_convec(int p, int nelem, int npoin, const int32_t *connec, const double *He,
                        const double *gpvol, const double *dlxigp_ip, const int32_t *invAtoIJK,
                        const int32_t *AtoI, const int32_t *AtoJ, const int32_t *AtoK,
                        const double *u, const double *q, const double *rho, const double *pr,
                        const double *E, double *Rmass, double *Rmom, double *Rener, int nthreads)
{
    const int Q = p + 1, nn = Q * Q * Q;
    const long NP = npoin;
    for (long i = 0; i < NP; ++i) { Rmass[i] = 0.0; Rener[i] = 0.0; }
    for (long i = 0; i < 3 * NP; ++i) Rmom[i] = 0.0;
#define INV(i, j, k) (invAtoIJK[((i) - 1) + Q * (((j) - 1) + Q * ((k) - 1))] - 1)
#define DL(g, d, n) dlxigp_ip[(g) + nn * ((d) + 3 * (n))]
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
#endif
    {
        long *ipoin = (long *)malloc(sizeof(long) * nn);
        double *buf = (double *)malloc(sizeof(double) * nn * 40);
        double *rhol = buf, *prl = buf + nn, *El = buf + 2 * nn, *REl = buf + 3 * nn,
               *kl = buf + 4 * nn, *Rkl = buf + 5 * nn, *ul = buf + 6 * nn, *ql = buf + 9 * nn,
               *fel = buf + 12 * nn, *fuel = buf + 15 * nn, *fkl = buf + 18 * nn,
               *fukl = buf + 21 * nn, *Re_mass = buf + 24 * nn, *Re_ener = buf + 25 * nn,
               *Re_mom = buf + 26 * nn;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (long e = 0; e < nelem; ++e) {
            /* gather (:230-252) */
            for (int a = 0; a < nn; ++a) {
                long ip = connec[e + (long)nelem * a] - 1;
                ipoin[a] = ip;
                kl[a] = 0.0;
                for (int d = 0; d < 3; ++d) kl[a] = kl[a] + u[ip + NP * d] * u[ip + NP * d] * 0.5;
                rhol[a] = rho[ip];
                El[a] = E[ip];
                REl[a] = rhol[a] * El[a];
                Rkl[a] = rhol[a] * kl[a];
                prl[a] = pr[ip];
                for (int d = 0; d < 3; ++d) {
                    ul[a + nn * d] = u[ip + NP * d];
                    ql[a + nn * d] = q[ip + NP * d];
                    /* energy-kernel products (:694-697) */
                    fel[a + nn * d] = rho[ip] * E[ip] * u[ip + NP * d];
                    fuel[a + nn * d] = E[ip] * u[ip + NP * d];
                    fkl[a + nn * d] = rho[ip] * kl[a] * u[ip + NP * d];
                    fukl[a + nn * d] = kl[a] * u[ip + NP * d];
                }
            }
            for (int g = 0; g < nn; ++g) {
                const int I = AtoI[g], J = AtoJ[g], K = AtoK[g];
                double gRho[3] = {0}, gP[3] = {0}, gE[3] = {0}, gRE[3] = {0}, gk[3] = {0}, gRk[3] = {0};
                double gU[3][3] = {{0}}, gQ[3][3] = {{0}}, gF[3][3][3] = {{{0}}}, gFuu[3][3][3] = {{{0}}};
                double gFe[3][3] = {{0}}, gFue[3][3] = {{0}}, gFk[3][3] = {{0}}, gFuk[3][3] = {{0}};
                for (int ii = 1; ii <= Q; ++ii) {
                    const int s[3] = {INV(ii, J, K), INV(I, ii, K), INV(I, J, ii)};
                    const double dl[3] = {DL(g, 0, ii - 1), DL(g, 1, ii - 1), DL(g, 2, ii - 1)};
                    for (int c = 0; c < 3; ++c) {
                        const int sc = s[c];
                        gRho[c] = gRho[c] + dl[c] * rhol[sc];
                        gP[c] = gP[c] + dl[c] * prl[sc];
                        gE[c] = gE[c] + dl[c] * (El[sc]);
                        gRE[c] = gRE[c] + dl[c] * (REl[sc]);
                        gk[c] = gk[c] + dl[c] * (kl[sc]);
                        gRk[c] = gRk[c] + dl[c] * (Rkl[sc]);
                    }
                    for (int id = 0; id < 3; ++id) {
                        for (int c = 0; c < 3; ++c) {
                            const int sc = s[c];
                            gU[id][c] = gU[id][c] + dl[c] * ul[sc + nn * id];
                            gQ[id][c] = gQ[id][c] + dl[c] * ql[sc + nn * id];
                            /* energy-kernel sweeps of precomputed products (:757-771) */
                            gFe[id][c] = gFe[id][c] + dl[c] * fel[sc + nn * id];
                            gFue[id][c] = gFue[id][c] + dl[c] * fuel[sc + nn * id];
                            gFk[id][c] = gFk[id][c] + dl[c] * fkl[sc + nn * id];
                            gFuk[id][c] = gFuk[id][c] + dl[c] * fukl[sc + nn * id];
                        }
                        for (int jd = 0; jd < 3; ++jd)
                            for (int c = 0; c < 3; ++c) {
                                const int sc = s[c];
                                /* fused-kernel products formed inside the sweep (:336-341) */
                                gF[id][jd][c] = gF[id][jd][c] + dl[c] * ql[sc + nn * id] * ul[sc + nn * jd];
                                gFuu[id][jd][c] = gFuu[id][jd][c] + dl[c] * ul[sc + nn * id] * ul[sc + nn * jd];
                            }
                    }
                }
                const double *H = He + 9 * ((long)g + (long)nn * e);
#define HE(i, j) H[(i) + 3 * (j)]
                double gradRho[3] = {0}, gradP[3] = {0}, gradE[3] = {0}, gradRE[3] = {0};
                double gradk[3] = {0}, gradRk[3] = {0}, gradU[3][3] = {{0}}, gradQ[3][3] = {{0}};
                double divF[3] = {0}, divFuu[3] = {0};
                double divFe = 0, divFue = 0, divFk = 0, divFuk = 0;
                for (int id = 0; id < 3; ++id)
                    for (int jd = 0; jd < 3; ++jd) {
                        gradRho[id] = gradRho[id] + HE(id, jd) * gRho[jd];
                        gradP[id] = gradP[id] + HE(id, jd) * gP[jd];
                        gradE[id] = gradE[id] + HE(id, jd) * gE[jd];
                        gradRE[id] = gradRE[id] + HE(id, jd) * gRE[jd];
                        divFe = divFe + HE(id, jd) * gFe[id][jd];
                        divFue = divFue + HE(id, jd) * gFue[id][jd];
                        gradk[id] = gradk[id] + HE(id, jd) * gk[jd];
                        gradRk[id] = gradRk[id] + HE(id, jd) * gRk[jd];
                        divFk = divFk + HE(id, jd) * gFk[id][jd];
                        divFuk = divFuk + HE(id, jd) * gFuk[id][jd];
                        for (int kd = 0; kd < 3; ++kd) {
                            gradU[id][jd] = gradU[id][jd] + HE(jd, kd) * gU[id][kd];
                            gradQ[id][jd] = gradQ[id][jd] + HE(jd, kd) * gQ[id][kd];
                            divF[id] = divF[id] + HE(jd, kd) * gF[id][jd][kd];
                            divFuu[id] = divFuu[id] + HE(jd, kd) * gFuu[id][jd][kd];
                        }
                    }
#undef HE
                const double divU = gradU[0][0] + gradU[1][1] + gradU[2][2];
                const double divQ = gradQ[0][0] + gradQ[1][1] + gradQ[2][2];
                double rm = 0.5 * (divQ + rhol[g] * divU);
                double re = 0.25 * (divFe + El[g] * divQ + REl[g] * divU + rhol[g] * divFue) +
                            0.25 * (divFk + kl[g] * divQ + Rkl[g] * divU + rhol[g] * divFuk);
                const double gv = gpvol[g + (long)nn * e];
                for (int id = 0; id < 3; ++id) {
                    double rmo = 0.25 * (divF[id] + ul[g + nn * id] * divQ + ql[g + nn * id] * divU +
                                         rhol[g] * divFuu[id]) + gradP[id];
                    rm = rm + 0.5 * (ul[g + nn * id] * gradRho[id]);
                    re = re + 0.25 * (ql[g + nn * id] * gradE[id] + ul[g + nn * id] * gradRE[id] +
                                      fuel[g + nn * id] * gradRho[id]) +
                         0.25 * (ql[g + nn * id] * gradk[id] + ul[g + nn * id] * gradRk[id] +
                                 fukl[g + nn * id] * gradRho[id]);
                    for (int jd = 0; jd < 3; ++jd) {
                        double aux = ul[g + nn * id] * ul[g + nn * jd];
                        rmo = rmo + 0.25 * (aux * gradRho[jd] + ql[g + nn * jd] * gradU[id][jd] +
                                            ul[g + nn * jd] * gradQ[id][jd]);
                    }
                    Re_mom[g + nn * id] = gv * rmo;
                }
                Re_mass[g] = gv * rm;
                Re_ener[g] = gv * re;
            }
            /* assembly (:412-426) */
            for (int a = 0; a < nn; ++a) {
                long ip = ipoin[a];
                ATOMIC_ADD(Rmass[ip], Re_mass[a]);
                ATOMIC_ADD(Rener[ip], Re_ener[a]);
                for (int d = 0; d < 3; ++d) ATOMIC_ADD(Rmom[ip + NP * d], Re_mom[a + nn * d]);
            }
        }
        free(buf);
        free(ipoin);
    }
}

/* ------------------------------------------------------------------------ */
/* Diffusion: kernels/sources/elem_diffu.f90:10-197                          */
/* ------------------------------------------------------------------------ */
void msod_oracle_diffu(int p, int nelem, int npoin, const int32_t *connec, const double *He,
                       const double *gpvol, const double *dlxigp_ip, const int32_t *invAtoIJK,
                       const int32_t *AtoI, const int32_t *AtoJ, const int32_t *AtoK, double Cp,
                       double Pr, double c_rho, double c_ener, const double *rho, const double *u,
                       const double *Tem, const double *mu_fluid, const double *mu_e,
                       const double *mu_sgs, double *Rmass, double *Rmom, double *Rener,
                       int nthreads)
{
    const int Q = p + 1, nn = Q * Q * Q;
    const long NP = npoin, NE = nelem;
    const double twoThirds = 2.0 / 3.0;
    for (long i = 0; i < NP; ++i) { Rmass[i] = 0.0; Rener[i] = 0.0; }
    for (long i = 0; i < 3 * NP; ++i) Rmom[i] = 0.0;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
#endif
    {
        long *ipoin = (long *)malloc(sizeof(long) * nn);
        double *buf = (double *)malloc(sizeof(double) * nn * 24);
        double *rhol = buf, *Teml = buf + nn, *mufl = buf + 2 * nn, *ul = buf + 3 * nn,
               *tauXl = buf + 6 * nn, *tauYl = buf + 9 * nn, *tauZl = buf + 12 * nn,
               *gradTl = buf + 15 * nn, *gradRhol = buf + 18 * nn, *tauUl = buf + 21 * nn;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (long e = 0; e < NE; ++e) {
            for (int a = 0; a < nn; ++a) {
                long ip = connec[e + NE * a] - 1;
                ipoin[a] = ip;
                rhol[a] = rho[ip];
                Teml[a] = Tem[ip];
                mufl[a] = mu_fluid[ip];
                for (int d = 0; d < 3; ++d) ul[a + nn * d] = u[ip + NP * d];
            }
            /* pass 1 (:67-139) */
            for (int g = 0; g < nn; ++g) {
                const double mu_fgp = mufl[g] + rhol[g] * mu_sgs[e + NE * g];
                const double mu_egp = mu_e[e + NE * g];
                const int I = AtoI[g], J = AtoJ[g], K = AtoK[g];
                double gIR[3] = {0}, gIT[3] = {0}, gIU[3][3] = {{0}};
                for (int ii = 1; ii <= Q; ++ii) {
                    const int s[3] = {INV(ii, J, K), INV(I, ii, K), INV(I, J, ii)};
                    for (int c = 0; c < 3; ++c) {
                        gIR[c] = gIR[c] + DL(g, c, ii - 1) * rhol[s[c]];
                        gIT[c] = gIT[c] + DL(g, c, ii - 1) * Teml[s[c]];
                    }
                    for (int id = 0; id < 3; ++id)
                        for (int c = 0; c < 3; ++c)
                            gIU[id][c] = gIU[id][c] + DL(g, c, ii - 1) * ul[s[c] + nn * id];
                }
                const double *H = He + 9 * ((long)g + (long)nn * e);
#define HE(i, j) H[(i) + 3 * (j)]
                double gradRho[3] = {0}, gradT[3] = {0}, gradU[3][3] = {{0}};
                for (int id = 0; id < 3; ++id)
                    for (int jd = 0; jd < 3; ++jd) {
                        gradRho[id] = gradRho[id] + HE(id, jd) * gIR[jd];
                        gradT[id] = gradT[id] + HE(id, jd) * gIT[jd];
                        for (int kd = 0; kd < 3; ++kd)
                            gradU[id][jd] = gradU[id][jd] + HE(jd, kd) * gIU[id][kd];
                    }
#undef HE
                const double divU = gradU[0][0] + gradU[1][1] + gradU[2][2];
                double tauU[3] = {0}, tau[3][3];
                for (int id = 0; id < 3; ++id) {
                    for (int jd = 0; jd < 3; ++jd) {
                        tauU[id] = tauU[id] + (mu_fgp + mu_egp) * (gradU[id][jd] + gradU[jd][id]) * ul[g + nn * jd];
                        tau[id][jd] = (mu_fgp + mu_egp) * (gradU[id][jd] + gradU[jd][id]);
                    }
                    tauU[id] = tauU[id] - (mu_fgp) * twoThirds * divU * ul[g + nn * id];
                    tau[id][id] = tau[id][id] - (mu_fgp) * twoThirds * divU;
                }
                for (int id = 0; id < 3; ++id) {
                    tauXl[g + nn * id] = tau[0][id];
                    tauYl[g + nn * id] = tau[1][id];
                    tauZl[g + nn * id] = tau[2][id];
                    gradTl[g + nn * id] = gradT[id];
                    gradRhol[g + nn * id] = gradRho[id];
                    tauUl[g + nn * id] = tauU[id];
                }
            }
            /* pass 2 (:141-193) */
            for (int g = 0; g < nn; ++g) {
                const double nu_e = c_rho * mu_e[e + NE * g] / rhol[g];
                const double kappa_e = mufl[g] * Cp / Pr + c_ener * mu_e[e + NE * g] * Cp / Pr +
                                       Cp * rhol[g] * mu_sgs[e + NE * g] / 0.9;
                const int I = AtoI[g], J = AtoJ[g], K = AtoK[g];
                const int tgt[3] = {I, J, K};
                double divDe = 0.0, divDr = 0.0, divDm[3] = {0};
                for (int ii = 1; ii <= Q; ++ii)
                    for (int id = 0; id < 3; ++id) {
                        const int s[3] = {INV(ii, J, K), INV(I, ii, K), INV(I, J, ii)};
                        for (int c = 0; c < 3; ++c) {
                            const int sc = s[c];
                            const double w = He[id + 3 * c + 9 * ((long)sc + (long)nn * e)] *
                                             gpvol[sc + (long)nn * e] * DL(sc, c, tgt[c] - 1);
                            divDr = divDr + w * gradRhol[sc + nn * id];
                        }
                        for (int c = 0; c < 3; ++c) {
                            const int sc = s[c];
                            const double w = He[id + 3 * c + 9 * ((long)sc + (long)nn * e)] *
                                             gpvol[sc + (long)nn * e] * DL(sc, c, tgt[c] - 1);
                            divDe = divDe + w * (tauUl[sc + nn * id] + kappa_e * gradTl[sc + nn * id]);
                        }
                        for (int c = 0; c < 3; ++c) {
                            const int sc = s[c];
                            const double w = He[id + 3 * c + 9 * ((long)sc + (long)nn * e)] *
                                             gpvol[sc + (long)nn * e] * DL(sc, c, tgt[c] - 1);
                            divDm[0] = divDm[0] + w * tauXl[sc + nn * id];
                        }
                        for (int c = 0; c < 3; ++c) {
                            const int sc = s[c];
                            const double w = He[id + 3 * c + 9 * ((long)sc + (long)nn * e)] *
                                             gpvol[sc + (long)nn * e] * DL(sc, c, tgt[c] - 1);
                            divDm[1] = divDm[1] + w * tauYl[sc + nn * id];
                        }
                        for (int c = 0; c < 3; ++c) {
                            const int sc = s[c];
                            const double w = He[id + 3 * c + 9 * ((long)sc + (long)nn * e)] *
                                             gpvol[sc + (long)nn * e] * DL(sc, c, tgt[c] - 1);
                            divDm[2] = divDm[2] + w * tauZl[sc + nn * id];
                        }
                    }
                const long ip = ipoin[g];
                ATOMIC_ADD(Rmass[ip], nu_e * divDr);
                ATOMIC_ADD(Rener[ip], divDe);
                for (int d = 0; d < 3; ++d) ATOMIC_ADD(Rmom[ip + NP * d], divDm[d]);
            }
        }
        free(buf);
        free(ipoin);
    }
#undef INV
#undef DL
}

/* ------------------------------------------------------------------------ */
/* Linear tetrahedra (TET04)                                                 */
/* ------------------------------------------------------------------------ */

/* The 4-point rule of quadrature_rules.f90:84-99 (the reference's literals)
 * and the P1 basis of elem_tet.f90:9-36 at its points, as the driver builds
 * them (main.f90:377-383): Ngp(g,:) = N(xgp(g,:)), dNgp(:,:,g) = dN. */
void msod_oracle_tet_tables(double *xgp, double *wgp, double *Ngp, double *dNgp)
{
    const double a = 0.5854101966249680, b = 0.1381966011250110;
    const double pts[4][3] = {{a, b, b}, {b, a, b}, {b, b, a}, {b, b, b}};
    for (int g = 0; g < 4; ++g) {
        for (int d = 0; d < 3; ++d) xgp[g + 4 * d] = pts[g][d];
        wgp[g] = 0.25;
        const double xi = pts[g][0], eta = pts[g][1], zeta = pts[g][2];
        Ngp[g + 4 * 0] = 1.0 - xi - eta - zeta;
        Ngp[g + 4 * 1] = xi;
        Ngp[g + 4 * 2] = eta;
        Ngp[g + 4 * 3] = zeta;
        for (int n = 0; n < 4; ++n)
            for (int d = 0; d < 3; ++d)
                dNgp[d + 3 * (n + 4 * g)] = n == 0 ? -1.0 : (d == n - 1 ? 1.0 : 0.0);
    }
}

/* jacobian_oper.f90:138-267, ndime = 3 branch */
void msod_oracle_tet_jacobian(int nelem, int npoin, const int32_t *connec, const double *coord,
                              const double *dNgp, const double *wgp, double *gpvol, double *He)
{
    for (int64_t e = 0; e < nelem; ++e) {
        for (int g = 0; g < 4; ++g) {
            double Je[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    double s = 0.0; /* dot_product(dNgp(i,:,g), coord(connec(e,:),j)) */
                    for (int n = 0; n < 4; ++n)
                        s += dNgp[i + 3 * (n + 4 * g)] * coord[(connec[e + (int64_t)nelem * n] - 1) + (int64_t)npoin * j];
                    Je[i][j] = 0.0 + s;
                }
            const double det = Je[0][0] * Je[1][1] * Je[2][2] + Je[0][1] * Je[1][2] * Je[2][0] +
                               Je[0][2] * Je[1][0] * Je[2][1] - Je[2][0] * Je[1][1] * Je[0][2] -
                               Je[2][1] * Je[1][2] * Je[0][0] - Je[2][2] * Je[1][0] * Je[0][1];
            double m[9];  /* a(1..9) of the reference, signs applied */
            m[0] = Je[1][1] * Je[2][2] - Je[2][1] * Je[1][2];
            m[1] = -(Je[1][0] * Je[2][2] - Je[2][0] * Je[1][2]);
            m[2] = Je[1][0] * Je[2][1] - Je[2][0] * Je[1][1];
            m[3] = -(Je[0][1] * Je[2][2] - Je[2][1] * Je[0][2]);
            m[4] = Je[0][0] * Je[2][2] - Je[2][0] * Je[0][2];
            m[5] = -(Je[0][0] * Je[2][1] - Je[2][0] * Je[0][1]);
            m[6] = Je[0][1] * Je[1][2] - Je[1][1] * Je[0][2];
            m[7] = -(Je[0][0] * Je[1][2] - Je[1][0] * Je[0][2]);
            m[8] = Je[0][0] * Je[1][1] - Je[1][0] * Je[0][1];
            /* b = transpose of a as 3x3 (b(2)=a(4), b(3)=a(7), ...), then /det */
            const int tr[9] = {0, 3, 6, 1, 4, 7, 2, 5, 8};
            double *h = He + 9 * (g + 4 * e);
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) /* He(r,c) = b(3r + c) */
                    h[r + 3 * c] = (1.0 / det) * m[tr[3 * r + c]];
            gpvol[g + 4 * e] = wgp[g] * det;
        }
    }
}

/* elem_convec.f90:839-948 */
void msod_oracle_fem_convec(int nelem, int npoin, const int32_t *connec, const double *Ngp,
                            const double *dNgp, const double *He, const double *gpvol,
                            const double *u, const double *q, const double *rho, const double *pr,
                            const double *E, double *Rmass, double *Rmom, double *Rener)
{
    const int64_t np = npoin, ne = nelem;
    memset(Rmass, 0, sizeof(double) * np);
    memset(Rmom, 0, sizeof(double) * np * 3);
    memset(Rener, 0, sizeof(double) * np);
    for (int64_t e = 0; e < ne; ++e) {
        int64_t ip[4];
        for (int n = 0; n < 4; ++n) ip[n] = connec[e + ne * n] - 1;
        double Re_mass[4] = {0}, Re_ener[4] = {0}, Re_mom[4][3] = {{0}};
        for (int g = 0; g < 4; ++g) {
            double divQ = 0, divFe = 0, gradRho[3] = {0}, gradE[3] = {0}, gradP[3] = {0};
            double gradU[3][3] = {{0}}, divF[3] = {0}, gpcar[4][3];
            const double *h = He + 9 * (g + 4 * e);
            for (int i = 0; i < 3; ++i)
                for (int n = 0; n < 4; ++n) {
                    const double *dn = dNgp + 3 * (n + 4 * g);
                    gpcar[n][i] = h[i] * dn[0] + h[i + 3] * dn[1] + h[i + 6] * dn[2];
                }
            for (int i = 0; i < 3; ++i)
                for (int n = 0; n < 4; ++n) {
                    const int64_t k = ip[n];
                    divQ = divQ + gpcar[n][i] * q[k + np * i];
                    divFe = divFe + gpcar[n][i] * (E[k] + pr[k]) * u[k + np * i];
                    gradRho[i] = gradRho[i] + gpcar[n][i] * rho[k];
                    gradE[i] = gradE[i] + gpcar[n][i] * (E[k] + pr[k]);
                    gradP[i] = gradP[i] + gpcar[n][i] * pr[k];
                    for (int j = 0; j < 3; ++j) {
                        gradU[i][j] = gradU[i][j] + gpcar[n][j] * u[k + np * i];
                        divF[i] = divF[i] + gpcar[n][j] * q[k + np * i] * u[k + np * j];
                    }
                }
            const double divU = gradU[0][0] + gradU[1][1] + gradU[2][2];
            for (int n = 0; n < 4; ++n) {
                const int64_t k = ip[n];
                double aux1 = 0, aux2 = 0;
                for (int i = 0; i < 3; ++i) {
                    aux1 = aux1 + gradRho[i] * u[k + np * i];
                    aux2 = aux2 + gradE[i] * u[k + np * i];
                }
                const double w = gpvol[g + 4 * e] * Ngp[g + 4 * n] * 0.5;
                Re_mass[n] = Re_mass[n] + w * (divQ + rho[k] * divU + aux1);
                Re_ener[n] = Re_ener[n] + w * (divFe + (E[k] + pr[k]) * divU + aux2);
                for (int i = 0; i < 3; ++i) {
                    double aux3 = 0, aux4 = 0;
                    for (int j = 0; j < 3; ++j) {
                        aux3 = aux3 + gradU[i][j] * q[k + np * j];
                        aux4 = aux4 + u[k + np * i] * u[k + np * j] * gradRho[j];
                    }
                    Re_mom[n][i] = Re_mom[n][i] + w * (divF[i] + q[k + np * i] * divU + aux3 + aux4 + gradP[i]);
                }
            }
        }
        for (int n = 0; n < 4; ++n) {
            Rmass[ip[n]] += Re_mass[n];
            Rener[ip[n]] += Re_ener[n];
            for (int i = 0; i < 3; ++i) Rmom[ip[n] + np * i] += Re_mom[n][i];
        }
    }
}

/* elem_diffu.f90:199-317.  As in the reference, the fluid properties at
 * point g weight node n with Ngp(n,g) (the table is stored (point, node)). */
void msod_oracle_fem_diffu(int nelem, int npoin, const int32_t *connec, const double *Ngp,
                           const double *dNgp, const double *He, const double *gpvol, double Cp,
                           double Pr, double c_rho, double c_ener, const double *rho,
                           const double *u, const double *Tem, const double *mu_fluid,
                           const double *mu_e, const double *mu_sgs, double *Rmass, double *Rmom,
                           double *Rener)
{
    const int64_t np = npoin, ne = nelem;
    memset(Rmass, 0, sizeof(double) * np);
    memset(Rmom, 0, sizeof(double) * np * 3);
    memset(Rener, 0, sizeof(double) * np);
    for (int64_t e = 0; e < ne; ++e) {
        int64_t ip[4];
        for (int n = 0; n < 4; ++n) ip[n] = connec[e + ne * n] - 1;
        double Re_mass[4] = {0}, Re_ener[4] = {0}, Re_mom[4][3] = {{0}};
        for (int g = 0; g < 4; ++g) {
            double auxRho = 0, auxNu = 0, auxMu = 0, auxKappa = 0;
            for (int n = 0; n < 4; ++n) auxRho = auxRho + Ngp[n + 4 * g] * rho[ip[n]];
            for (int n = 0; n < 4; ++n) {
                const double me = mu_e[e + ne * n], ms = mu_sgs[e + ne * n], mf = mu_fluid[ip[n]];
                auxNu = auxNu + Ngp[n + 4 * g] * me * (c_rho / auxRho);
                auxMu = auxMu + Ngp[n + 4 * g] * (mf + auxRho * ms + me);
                auxKappa = auxKappa + Ngp[n + 4 * g] * (mf * (Cp / Pr) + c_ener * me / 0.4 + auxRho * ms / 0.9);
            }
            double gpcar[4][3];
            const double *h = He + 9 * (g + 4 * e);
            for (int i = 0; i < 3; ++i)
                for (int n = 0; n < 4; ++n) {
                    const double *dn = dNgp + 3 * (n + 4 * g);
                    gpcar[n][i] = h[i] * dn[0] + h[i + 3] * dn[1] + h[i + 6] * dn[2];
                }
            double gradRho[3] = {0}, gradT[3] = {0}, gradU[3][3] = {{0}};
            for (int i = 0; i < 3; ++i)
                for (int n = 0; n < 4; ++n) {
                    const int64_t k = ip[n];
                    gradRho[i] = gradRho[i] + gpcar[n][i] * rho[k];
                    gradT[i] = gradT[i] + gpcar[n][i] * Tem[k];
                    for (int j = 0; j < 3; ++j) gradU[i][j] = gradU[i][j] + gpcar[n][i] * u[k + np * j];
                }
            const double divU = gradU[0][0] + gradU[1][1] + gradU[2][2];
            double tau[3][3] = {{0}}, tauU[3] = {0};
            for (int i = 0; i < 3; ++i) {
                tau[i][i] = -(2.0 / 3.0) * divU;
                for (int j = 0; j < 3; ++j) tau[i][j] = tau[i][j] + gradU[i][j] + gradU[j][i];
            }
            for (int j = 0; j < 3; ++j) {
                double auxU = 0;
                for (int n = 0; n < 4; ++n) auxU = auxU + Ngp[n + 4 * g] * u[ip[n] + np * j];
                for (int i = 0; i < 3; ++i) tauU[i] = tauU[i] + tau[i][j] * auxU;
            }
            const double gv = gpvol[g + 4 * e];
            for (int i = 0; i < 3; ++i)
                for (int n = 0; n < 4; ++n) {
                    Re_mass[n] = Re_mass[n] + gv * auxNu * gpcar[n][i] * gradRho[i];
                    Re_ener[n] = Re_ener[n] + gv * gpcar[n][i] * (auxMu * tauU[i] + auxKappa * gradT[i]);
                    for (int j = 0; j < 3; ++j)
                        Re_mom[n][i] = Re_mom[n][i] + gv * gpcar[n][j] * auxMu * tau[i][j];
                }
        }
        for (int n = 0; n < 4; ++n) {
            Rmass[ip[n]] += Re_mass[n];
            Rener[ip[n]] += Re_ener[n];
            for (int i = 0; i < 3; ++i) Rmom[ip[n] + np * i] += Re_mom[n][i];
        }
    }
}
