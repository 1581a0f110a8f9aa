/*
 * oracle.c -- CPU restatement of fem25's hot path (TEST INFRASTRUCTURE ONLY).
 *
 * This file is the parity checker for the HIP implementation in fem25_amd/.
 * It is loaded by tests/, by __graft_entry__.smoke() and by the cpu_baseline
 * leg of bench.py, never by the product.  Every function follows the loop
 * structure of the reference routine it restates (cited per function), so the
 * floating-point work happens in the same order wherever that is practical.
 *
 * How each function is pinned (checks live in tests/):
 *   orc_gll          numpy Legendre closed form (roots of P'_{n-1}, weights
 *                    2/(n(n-1)P_{n-1}^2)); compiled reference GAUSSQ tables.
 *   orc_dshl         cardinal property at the GLL nodes, partition of unity,
 *                    exact derivative of polynomials of degree <= p; compiled
 *                    reference DSHL tables (p = 4, 5, 6).
 *   orc_csf2d        closed form + partition of unity.
 *   orc_cjmk_tvi     affine rectangles: det J = hx*hz/4, diagonal inverse;
 *                    compiled reference CJMK_TVI.
 *   orc_km25isoq     rigid translations in the null space of K0, symmetry /
 *                    Hermiticity, mass-block sum = -w^2 int rho + ky^2 int C;
 *                    compiled reference KM25ISOQ.
 *   orc_zdampfn,
 *   orc_cjaccq,
 *   orc_cm25isoq     reduces to KM25ISOQ when the stretch factor is 1;
 *                    compiled reference CJACCQ/CM25ISOQ on PML elements.
 *   orc_genien..     brute-force adjacency pattern (GENCRS_DUMB idea),
 *   orc_gencrs       node-count formulas of the structured mesh; compiled
 *                    reference GENIEN/GENID/GENLM/GENCRS_QUICK where it builds.
 *   orc_asmble_*     element-sum identity against KM25ISOQ, quadratic-in-ky /
 *                    affine-in-w^2 identity; compiled reference ASMBLE_DRIVER.
 *   orc_djacab,
 *   orc_dkm_vpvs     linearity in the collocation field; compiled reference.
 *   orc_cfmat_dist   compiled reference CFMAT_DIST/VFORCE25.
 *   orc_adjflam_dist hand-computed CSC gather on a 3-column example.
 *   orc_pchadj_jac   dense J = G^T F and diag(J^H J) by numpy (pchadj.f90
 *                    cannot be compiled here: it needs MUMPS headers).
 *   orc_gnmfree      compiled reference GNMFREE reading Jacobian files in its
 *                    own format; dense Re(J^H J p) by numpy.
 *   orc_grad_exact   central finite differences of the objective (exact for a
 *                    quadratic) on small meshes.
 *   orc_djacab2,
 *   orc_dabmkisob25,
 *   orc_dabm2kisob25 compiled reference DJACAB2/DABMKISOB25/DABM2KISOB25.
 *   orc_qp_nu_alpha  RHSALP and RHSBNU: summed over every column and element
 *                    they are S[lambda = l_L, mu = l_M, rho = 0] u and its
 *                    conjugate transpose on lambda, S assembled by
 *                    orc_asmble_z (pinned above) from the anchor field
 *                    l = 2 Vp V (PP) or (-4 Vs V, 2 Vs V) (SS); P3 on anchors
 *                    with one element, where the same identity fixes the
 *                    column; elsewhere the column is the last element's, the
 *                    element choice CFMAT_DIST makes (compiled reference).
 *   orc_met225       as P3 (qpmfree.f90 needs MUMPS headers: not compiled).
 *   orc_exresp25d,
 *   orc_crotate,
 *   orc_cphm2cm,
 *   orc_cresid,
 *   orc_bpresid4     compiled reference EXRESP25D, CROTATE, CPHM2CM, CRESID,
 *                    BPRESID4 (every norm, residual type, phase gate).
 *   orc_hdiffer,
 *   orc_resid_drhs   closed forms (sqrt(-i w) = sqrt(w/2)(1 - i), ...); the
 *                    scatter by brute force with repeated and invalid DOFs
 *                    (adjgrad.f90 needs MUMPS headers: not compiled).
 *   orc_cadjlam_rhs  composition of the above; with equal weights on the two
 *                    horizontal components the L2 residual rotated back is
 *                    W^2 (R(-az) OBS - EST) (closed form).
 *   orc_rec_space1d,
 *   orc_laplace_1dv2,
 *   orc_perm_obs,
 *   orc_bpresid_covd4 compiled reference REC_SPACE1D, LAPLACE_1DV2, PERM_OBS,
 *                    BPRESID_COVD4; with every receiver observed and unit
 *                    weights the residual is the second difference of the
 *                    residual along the line (closed form), rows sum to zero
 *                    and <WORK, RESIDV> = sum (r_k+1 - r_k)^2 / h_k.
 *   orc_fill_uwave   compiled reference FILL_UWAVE_QMESH; closed form of the
 *                    plane wave at the GLL nodes.
 *   orc_cadjlam_rhs_covd composition of the above (compiled LAPLACE_1DV2,
 *                    PERM_OBS and BPRESID_COVD4 behind the same frames).
 */
#include "oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define NDIM 3

/* ------------------------------------------------------------------------ */
/* gaussq.f:5-130 (GAUSSQ with kind=5, alpha=beta=0, kpts=2, endpoints -1, 1):
 * the Golub-Welsch construction of the Gauss-Lobatto-Legendre rule --
 * Jacobi recurrence coefficients (class, gaussq.f:160-262), Lobatto
 * modification of the last row (solve, gaussq.f:132-155; gaussq.f:96-101),
 * implicit-QL eigenvalues with first eigenvector components (gausq2,
 * gaussq.f:266-403).  Restated step by step: the nodes' last bits decide
 * GENID's fixed-DOF test (graph25.f:437-469) and must match. */
static double gq_solve(double shift, int n, const double *a, const double *b)
{
    double alpha = a[1] - shift;
    for (int i = 2; i <= n - 1; i++) alpha = a[i] - shift - (b[i - 1] * b[i - 1]) / alpha;
    return 1.0 / alpha;
}

static void gq_gausq2(int n, double *d, double *e, double *z)
{
    const double machep = 2.220446049250313e-16;
    if (n == 1) return;
    e[n] = 0.0;
    for (int l = 1; l <= n; l++) {
        int j = 0, m;
        for (;;) {
            for (m = l; m <= n; m++) {
                if (m == n) break;
                if (fabs(e[m]) <= machep * (fabs(d[m]) + fabs(d[m + 1]))) break;
            }
            double p = d[l];
            if (m == l) break;
            if (j == 30) return;
            j++;
            double g = (d[l + 1] - p) / (2.0 * e[l]);
            double r = sqrt(g * g + 1.0);
            g = d[m] - p + e[l] / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
            double s = 1.0, c = 1.0;
            p = 0.0;
            for (int ii = 1; ii <= m - l; ii++) {
                int i = m - ii;
                double f = s * e[i], b = c * e[i];
                if (fabs(f) < fabs(g)) {
                    s = f / g; r = sqrt(s * s + 1.0); e[i + 1] = g * r; c = 1.0 / r; s = s * c;
                } else {
                    c = g / f; r = sqrt(c * c + 1.0); e[i + 1] = f * r; s = 1.0 / r; c = c * s;
                }
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                f = z[i + 1];
                z[i + 1] = s * z[i] + c * f;
                z[i] = c * z[i] - s * f;
            }
            d[l] = d[l] - p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    for (int ii = 2; ii <= n; ii++) {
        int i = ii - 1, k = i;
        double p = d[i];
        for (int jj = ii; jj <= n; jj++)
            if (d[jj] < p) { k = jj; p = d[jj]; }
        if (k != i) { d[k] = d[i]; d[i] = p; p = z[i]; z[i] = z[k]; z[k] = p; }
    }
}

void orc_gll(int n, double *x, double *w)
{
    double t[66], b[66], z[66];
    /* class, kind = 5 with alpha = beta = 0 */
    double ab = 0.0, abi = 2.0 + ab, a2b2 = 0.0;
    double muzero = 2.0;          /* 2^(ab+1) gamma(1) gamma(1) / gamma(2) */
    t[1] = 0.0 / abi;
    b[1] = sqrt(4.0 * (1.0 + 0.0) * (1.0 + 0.0) / ((abi + 1.0) * abi * abi));
    for (int i = 2; i <= n - 1; i++) {
        abi = 2.0 * i + ab;
        t[i] = a2b2 / ((abi - 2.0) * abi);
        b[i] = sqrt(4.0 * i * (i + 0.0) * (i + 0.0) * (i + ab) / ((abi * abi - 1.0) * abi * abi));
    }
    abi = 2.0 * n + ab;
    t[n] = a2b2 / ((abi - 2.0) * abi);
    /* Lobatto: both end points fixed (gaussq.f:96-101) */
    double gam = gq_solve(-1.0, n, t, b);
    double t1 = (-1.0 - 1.0) / (gq_solve(1.0, n, t, b) - gam);
    b[n - 1] = sqrt(t1);
    t[n] = -1.0 + gam * t1;
    z[1] = 1.0;
    for (int i = 2; i <= n; i++) z[i] = 0.0;
    gq_gausq2(n, t, b, z);
    for (int i = 1; i <= n; i++) {
        x[i - 1] = t[i];
        w[i - 1] = muzero * z[i] * z[i];
    }
}

/* dshl.f:266-314 POLCOE: monomial coefficients of the interpolant of y on x */
static void polcoe(const double *x, const double *y, int n, double *cof, double *s)
{
    for (int i = 0; i < n; i++) { s[i] = 0.0; cof[i] = 0.0; }
    s[n - 1] = -x[0];
    for (int i = 1; i < n; i++) {
        for (int j = n - 1 - i; j <= n - 2; j++) s[j] -= x[i] * s[j + 1];
        s[n - 1] -= x[i];
    }
    for (int j = 0; j < n; j++) {
        double phi = (double)n;
        for (int k = n - 1; k >= 1; k--) phi = k * s[k] + x[j] * phi;
        double ff = y[j] / phi, b = 1.0;
        for (int k = n; k >= 1; k--) {
            cof[k - 1] += b * ff;
            b = s[k - 1] + x[j] * b;
        }
    }
}

static double ipow(double x, int k)
{
    double r = 1.0;
    for (int i = 0; i < k; i++) r *= x;
    return r;
}

/* dshl.f:247-261 SHAPEF */
static double shapef(int n, double xi, const double *cof)
{
    double s = 0.0;
    for (int i = n; i >= 2; i--) s += cof[i - 1] * ipow(xi, i - 1);
    s += cof[0];
    if (fabs(s) < 1.11e-15) s = 0.0;
    return s;
}

/* dshl.f:1-243 DSHL with IITYPE=1 (GLL interpolation points) */
void orc_dshl(int nl, int nint, double *xipts, double *xigll, double *wgll, double *shl)
{
    const double eps = 2.11e-16;
    double wt[64];
    int nen = nl * nl;
    orc_gll(nl, xipts, wt);
    if (nl % 2) xipts[(nl - 1) / 2] = 0.0;
    orc_gll(nint, xigll, wgll);
    if (nint % 2) xigll[(nint - 1) / 2] = 0.0;
    double lagr[64], cofxi[64], cofeta[64], s[64], xil[64], etal[64];
    int ia = 0;
    for (int ieta = 0; ieta < nl; ieta++) {
        for (int k = 0; k < nl; k++) lagr[k] = (k == ieta) ? 1.0 : 0.0;
        polcoe(xipts, lagr, nl, cofeta, s);
        for (int iz = 0; iz < nint; iz++) {
            double sumd = 0.0;
            for (int k = nl; k >= 2; k--) sumd += (k - 1.0) * cofeta[k - 1] * ipow(xigll[iz], k - 2);
            if (fabs(sumd) < eps) sumd = 0.0;
            etal[iz] = sumd;
        }
        for (int ixi = 0; ixi < nl; ixi++, ia++) {
            for (int k = 0; k < nl; k++) lagr[k] = (k == ixi) ? 1.0 : 0.0;
            polcoe(xipts, lagr, nl, cofxi, s);
            for (int ix = 0; ix < nint; ix++) {
                double sumd = 0.0;
                for (int k = nl; k >= 2; k--) sumd += (k - 1.0) * cofxi[k - 1] * ipow(xigll[ix], k - 2);
                if (fabs(sumd) < eps) sumd = 0.0;
                xil[ix] = sumd;
            }
            for (int iz = 0; iz < nint; iz++) {
                for (int ix = 0; ix < nint; ix++) {
                    double *o = shl + 3 * (ia + nen * (ix + nint * iz));
                    double sna = shapef(nl, xigll[iz], cofeta);
                    double snb = shapef(nl, xigll[ix], cofxi);
                    o[2] = sna * snb;
                    if (fabs(o[2]) < eps) o[2] = 0.0;
                    o[0] = sna * xil[ix];
                    o[1] = snb * etal[iz];
                    if (fabs(o[0]) < eps) o[0] = 0.0;
                    if (fabs(o[1]) < eps) o[1] = 0.0;
                }
            }
        }
    }
}

/* csf2d.f:1-131 CSF2D for NGNOD=4 */
int orc_csf2d(double xi, double eta, double *sf, double *dsf)
{
    const double eps = 2.2204e-16;
    sf[0] = .25 * (1.0 - xi) * (1.0 - eta);
    sf[1] = .25 * (1.0 + xi) * (1.0 - eta);
    sf[2] = .25 * (1.0 + xi) * (1.0 + eta);
    sf[3] = .25 * (1.0 - xi) * (1.0 + eta);
    dsf[0 + 2 * 0] = .25 * (-1.0) * (1.0 - eta);
    dsf[0 + 2 * 1] = .25 * (1.0) * (1.0 - eta);
    dsf[0 + 2 * 2] = .25 * (1.0) * (1.0 + eta);
    dsf[0 + 2 * 3] = .25 * (-1.0) * (1.0 + eta);
    dsf[1 + 2 * 0] = .25 * (1.0 - xi) * (-1.0);
    dsf[1 + 2 * 1] = .25 * (1.0 + xi) * (-1.0);
    dsf[1 + 2 * 2] = .25 * (1.0 + xi) * (1.0);
    dsf[1 + 2 * 3] = .25 * (1.0 - xi) * (1.0);
    double s2 = 0, s3 = 0;
    for (int i = 0; i < 4; i++) {
        if (fabs(sf[i]) < eps) sf[i] = 0.0;
        if (fabs(dsf[2 * i]) < eps) dsf[2 * i] = 0.0;
        if (fabs(dsf[2 * i + 1]) < eps) dsf[2 * i + 1] = 0.0;
        s2 += dsf[2 * i];
        s3 += dsf[2 * i + 1];
    }
    if (fabs(s2) > eps || fabs(s3) > eps) return 1;
    return 0;
}

/* csf2d.f:135-200 CSF2DND (NGNOD=4) */
static void csf2dnd(double xi, double eta, double *sf)
{
    const double eps = 2.2204e-16;
    sf[0] = .25 * (1.0 - xi) * (1.0 - eta);
    sf[1] = .25 * (1.0 + xi) * (1.0 - eta);
    sf[2] = .25 * (1.0 + xi) * (1.0 + eta);
    sf[3] = .25 * (1.0 - xi) * (1.0 + eta);
    for (int i = 0; i < 4; i++)
        if (fabs(sf[i]) < eps) sf[i] = 0.0;
}

/* shared geometric part of cjac.f:61-135 and djac.f:73-135 */
static int jacobian_at(double xi, double eta, const int *ieng_e, const double *xlocs,
                       const double *zlocs, double *sf, double *rjac, double *dxdxi,
                       double *dxdeta, double *dzdxi, double *dzdeta)
{
    double dsf[8];
    if (orc_csf2d(xi, eta, sf, dsf)) return 1;
    double a = 0, b = 0, c = 0, d = 0;
    for (int k = 0; k < 4; k++) {
        int g = ieng_e[k] - 1;
        a += dsf[2 * k] * xlocs[g];
        b += dsf[2 * k + 1] * xlocs[g];
        c += dsf[2 * k] * zlocs[g];
        d += dsf[2 * k + 1] * zlocs[g];
    }
    *dxdxi = a; *dxdeta = b; *dzdxi = c; *dzdeta = d;
    *rjac = a * d - b * c;
    if (fabs(*rjac) <= 2.22e-16 || *rjac < 0.0) return 2;
    return 0;
}

static void fill_shg(int nen, const double *shl_q, double *shg_q, double rjac, double dxdxi,
                     double dxdeta, double dzdxi, double dzdeta)
{
    for (int a = 0; a < nen; a++) {
        double drnxi = shl_q[3 * a], drneta = shl_q[3 * a + 1];
        shg_q[3 * a + 2] = shl_q[3 * a + 2];
        shg_q[3 * a + 0] = (drnxi * dzdeta - drneta * dzdxi) / rjac;
        shg_q[3 * a + 1] = -(drnxi * dxdeta - drneta * dxdxi) / rjac;
    }
}

/* cjac.f:1-138 CJMK_TVI, isotropic (NCOEFF=2) */
int orc_cjmk_tvi(int nen, int nint, const int *ieng_e, const double *xigll,
                 const double *xlocs, const double *zlocs, const double *dens,
                 const double *ecoeff, int nnpg, const double *shl,
                 double *det, double *rho, double *dmat, double *shg)
{
    for (int ix = 0; ix < nint; ix++) {
        for (int iz = 0; iz < nint; iz++) {
            double sf[4], rjac, a, b, c, d;
            int err = jacobian_at(xigll[ix], xigll[iz], ieng_e, xlocs, zlocs, sf, &rjac, &a, &b, &c, &d);
            int q = ix + nint * iz;
            double r = 0, l = 0, m = 0;
            for (int k = 0; k < 4; k++) {
                int g = ieng_e[k] - 1;
                l += sf[k] * ecoeff[g];
                m += sf[k] * ecoeff[g + nnpg];
                r += sf[k] * dens[g];
            }
            dmat[q] = l; dmat[q + nint * nint] = m; rho[q] = r;
            det[q] = rjac;
            if (err) return 1;
            fill_shg(nen, shl + 3 * nen * q, shg + 3 * nen * q, rjac, a, b, c, d);
        }
    }
    return 0;
}

/* cjac.f:818-850 ZDAMPFN */
double complex orc_zdampfn(double rn, double vp, double delta, double dmax_in,
                           double rcoeff, double rkappa, double fcent, double omega)
{
    double dmax = (dmax_in > 0.0) ? dmax_in : -3.0 * vp * log(rcoeff) / (2.0 * delta);
    double d = dmax * (rn / delta) * (rn / delta);
    double alpha = M_PI * fcent * (1.0 - rn / delta);
    double complex z = rkappa + fabs(d) / (alpha + I * omega);
    return 1.0 / z;
}

/* cjac.f:399-581 CJACCQ, isotropic */
int orc_cjaccq(int nen, int nint, double omega, double xwidth, double zwidth,
               double rcoeff, double dmax_in, double rkappa, double fcent,
               const int *ieng_e, const double *xigll, const double *xd, const double *zd,
               const double *xlocs, const double *zlocs, const double *dens,
               const double *ecoeff, int nnpg, const double *shl,
               double *det, double *rho, double *dmat, double *shg,
               double complex *dampx, double complex *dampz)
{
    const double eps = 2.22e-16;
    for (int ix = 0; ix < nint; ix++) {
        for (int iz = 0; iz < nint; iz++) {
            double sf[4], rjac, a, b, c, d;
            int err = jacobian_at(xigll[ix], xigll[iz], ieng_e, xlocs, zlocs, sf, &rjac, &a, &b, &c, &d);
            int q = ix + nint * iz;
            double r = 0, l = 0, m = 0, xdep = 0, zdep = 0, vel = 0;
            for (int k = 0; k < 4; k++) {
                int g = ieng_e[k] - 1;
                l += sf[k] * ecoeff[g];
                m += sf[k] * ecoeff[g + nnpg];
                r += sf[k] * dens[g];
                xdep += sf[k] * xd[g];
                zdep += sf[k] * zd[g];
                vel += sf[k] * sqrt((ecoeff[g] + 2.0 * ecoeff[g + nnpg]) / dens[g]);
            }
            dmat[q] = l; dmat[q + nint * nint] = m; rho[q] = r;
            dampx[q] = (xdep > eps) ? orc_zdampfn(xdep, vel, xwidth, dmax_in, rcoeff, rkappa, fcent, omega) : 1.0;
            dampz[q] = (zdep > eps) ? orc_zdampfn(zdep, vel, zwidth, dmax_in, rcoeff, rkappa, fcent, omega) : 1.0;
            det[q] = rjac;
            if (err) return 1;
            fill_shg(nen, shl + 3 * nen * q, shg + 3 * nen * q, rjac, a, b, c, d);
        }
    }
    return 0;
}

/* estiff25.f:1-164 KM25ISOQ */
void orc_km25isoq(int nen, int nint, double omega, double py, const double *wts,
                  const double *vlam, const double *vmu, const double *rho,
                  const double *det, const double *shg, double *me, double complex *ke)
{
    int nee = NDIM * nen;
    for (int i = 0; i < nee * nee; i++) { me[i] = 0.0; ke[i] = 0.0; }
    double wpy = omega * py, py2 = py * py, w2 = omega * omega;
    for (int ix = 0; ix < nint; ix++) {
        for (int iz = 0; iz < nint; iz++) {
            int q = ix + nint * iz;
            double weight = wts[ix] * wts[iz] * det[q];
            double lam = vlam[q] * weight, mu = vmu[q] * weight;
            double c11 = lam + 2.0 * mu, c12 = lam, c13 = lam, c22 = lam + 2.0 * mu, c23 = lam;
            double c33 = lam + 2.0 * mu, c44 = mu, c55 = mu, c66 = mu;
            double rhow = rho[q] * weight;
            double cm1 = -w2 * (rhow - py2 * c66), cm2 = -w2 * (rhow - py2 * c22), cm3 = -w2 * (rhow - py2 * c44);
            const double *sq = shg + 3 * nen * q;
            for (int ib = 0; ib < nen; ib++) {
                double phib = sq[3 * ib + 2], phibx = sq[3 * ib], phibz = sq[3 * ib + 1];
                double r1 = phib * cm1, r2 = phib * cm2, r3 = phib * cm3;
                double k11t1 = c11 * phibx, k11t2 = c55 * phibz, k13t1 = c13 * phibz, k13t2 = c55 * phibx;
                double k22t1 = c66 * phibx, k22t2 = c44 * phibz, k31t1 = c13 * phibx, k31t2 = c55 * phibz;
                double k33t1 = c33 * phibz, k33t2 = c55 * phibx;
                double k12t1 = wpy * c12 * phib, k12t2 = wpy * c66 * phibx;
                double k21t1 = wpy * c12 * phibx, k21t2 = wpy * c66 * phib;
                double k23t1 = wpy * c23 * phibz, k23t2 = wpy * c44 * phib;
                double k32t1 = wpy * c23 * phib, k32t2 = wpy * c44 * phibz;
                for (int ia = 0; ia < nen; ia++) {
                    double phia = sq[3 * ia + 2], phiax = sq[3 * ia], phiaz = sq[3 * ia + 1];
                    int ip = NDIM * ia, iq = NDIM * ib;
                    me[ip + nee * iq] += r1 * phia;
                    me[ip + 1 + nee * (iq + 1)] += r2 * phia;
                    me[ip + 2 + nee * (iq + 2)] += r3 * phia;
                    double ke11 = phiax * k11t1 + phiaz * k11t2;
                    double ke13 = phiax * k13t1 + phiaz * k13t2;
                    double ke22 = phiax * k22t1 + phiaz * k22t2;
                    double ke31 = phiaz * k31t1 + phiax * k31t2;
                    double ke33 = phiaz * k33t1 + phiax * k33t2;
                    double ke12 = phiax * k12t1 - phia * k12t2;
                    double ke21 = phia * k21t1 - phiax * k21t2;
                    double ke23 = phia * k23t1 - phiaz * k23t2;
                    double ke32 = phiaz * k32t1 - phia * k32t2;
                    ke[ip + nee * iq] += ke11;
                    ke[ip + nee * (iq + 1)] -= I * ke12;
                    ke[ip + nee * (iq + 2)] += ke13;
                    ke[ip + 1 + nee * iq] += I * ke21;
                    ke[ip + 1 + nee * (iq + 1)] += ke22;
                    ke[ip + 1 + nee * (iq + 2)] += I * ke23;
                    ke[ip + 2 + nee * iq] += ke31;
                    ke[ip + 2 + nee * (iq + 1)] -= I * ke32;
                    ke[ip + 2 + nee * (iq + 2)] += ke33;
                }
            }
        }
    }
}

/* estiff25.f:168-334 CM25ISOQ */
void orc_cm25isoq(int nen, int nint, double omega, double py, const double *wts,
                  const double *vlam, const double *vmu, const double *rho,
                  const double *det, const double *shg, const double complex *dampx,
                  const double complex *dampz, double *me, double complex *ce)
{
    int nee = NDIM * nen;
    for (int i = 0; i < nee * nee; i++) { me[i] = 0.0; ce[i] = 0.0; }
    double wpy = omega * py, py2 = py * py, w2 = omega * omega;
    double complex cwpy = I * wpy;
    for (int ix = 0; ix < nint; ix++) {
        for (int iz = 0; iz < nint; iz++) {
            int q = ix + nint * iz;
            double weight = wts[ix] * wts[iz] * det[q];
            double complex ex = dampx[q], ez = dampz[q];
            double lam = vlam[q] * weight, mu = vmu[q] * weight;
            double complex clam = lam, cmu = mu;
            double complex c11 = clam + cmu + cmu, c12 = clam, c13 = clam, c23 = clam;
            double complex c33 = clam + cmu + cmu, c44 = cmu, c55 = cmu, c66 = cmu;
            double rhow = rho[q] * weight;
            double cm1 = -w2 * (rhow - py2 * mu), cm2 = -w2 * (rhow - py2 * (lam + 2.0 * mu)), cm3 = -w2 * (rhow - py2 * mu);
            const double *sq = shg + 3 * nen * q;
            for (int ib = 0; ib < nen; ib++) {
                double phib = sq[3 * ib + 2];
                double complex phibx = sq[3 * ib] * ex, phibz = sq[3 * ib + 1] * ez;
                double r1 = cm1 * phib, r2 = cm2 * phib, r3 = cm3 * phib;
                double complex k11t1 = c11 * phibx, k11t2 = c55 * phibz, k13t1 = c13 * phibz, k13t2 = c55 * phibx;
                double complex k22t1 = c66 * phibx, k22t2 = c44 * phibz, k31t1 = c13 * phibx, k31t2 = c55 * phibz;
                double complex k33t1 = c33 * phibz, k33t2 = c55 * phibx;
                double complex k12t1 = cwpy * c12 * phib, k12t2 = cwpy * c66 * phibx;
                double complex k21t1 = cwpy * c12 * phibx, k21t2 = cwpy * c66 * phib;
                double complex k23t1 = cwpy * c23 * phibz, k23t2 = cwpy * c44 * phib;
                double complex k32t1 = cwpy * c23 * phib, k32t2 = cwpy * c44 * phibz;
                for (int ia = 0; ia < nen; ia++) {
                    double phia = sq[3 * ia + 2];
                    double complex phiax = sq[3 * ia] * ex, phiaz = sq[3 * ia + 1] * ez;
                    int ip = NDIM * ia, iq = NDIM * ib;
                    me[ip + nee * iq] += phia * r1;
                    me[ip + 1 + nee * (iq + 1)] += phia * r2;
                    me[ip + 2 + nee * (iq + 2)] += phia * r3;
                    double complex ce11 = phiax * k11t1 + phiaz * k11t2;
                    double complex ce13 = phiax * k13t1 + phiaz * k13t2;
                    double complex ce22 = phiax * k22t1 + phiaz * k22t2;
                    double complex ce31 = phiaz * k31t1 + phiax * k31t2;
                    double complex ce33 = phiaz * k33t1 + phiax * k33t2;
                    double complex ce12 = phiax * k12t1 - phia * k12t2;
                    double complex ce21 = phia * k21t1 - phiax * k21t2;
                    double complex ce23 = phia * k23t1 - phiaz * k23t2;
                    double complex ce32 = phiaz * k32t1 - phia * k32t2;
                    ce[ip + nee * iq] += ce11;
                    ce[ip + nee * (iq + 1)] -= ce12;
                    ce[ip + nee * (iq + 2)] += ce13;
                    ce[ip + 1 + nee * iq] += ce21;
                    ce[ip + 1 + nee * (iq + 1)] += ce22;
                    ce[ip + 1 + nee * (iq + 2)] += ce23;
                    ce[ip + 2 + nee * iq] += ce31;
                    ce[ip + 2 + nee * (iq + 1)] -= ce32;
                    ce[ip + 2 + nee * (iq + 2)] += ce33;
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* haskattn.f:602-641 VCMPLX: complex velocity of a dispersive medium (Aki &
 * Richards 5.94): V (1 + ln(f/f0)/(pi Q)) - i V/(2Q), the real factor
 * clamped below at 0.1; real V when FREQ0 = 0, Q = 9999 or Q = 0. */
double complex orc_vcmplx(double freq, double freq0, double v, double q)
{
    const double pi = 3.1415926535897932384626434;
    if (freq0 == 0.0 || fabs(q - 9999.0) < 1e-5) return v;
    if (q == 0.0) return v;
    double fac1 = log(freq / freq0);
    double fact = 1.0 + 1.0 / (pi * q) * fac1;
    if (fact < 0.1) fact = 0.1;
    return v * fact + I * (-v / (2.0 * q));
}

/* cjac.f:142-250 CJKM_QUAL (and the geometry part of CJCM_QUAL,
 * cjac.f:585-720): VP, VS, QP, QS interpolated bilinearly from the anchors
 * (velocities from ECOEFF and DENS at each anchor), DET, SHG; with xd/zd
 * non-NULL also DAMPX/DAMPZ from ZDAMPFN with VEL = VP. */
int orc_cjkm_qual(int nen, int nint, const int *ieng_e, const double *xigll,
                  const double *xlocs, const double *zlocs, const double *dens,
                  const double *qp, const double *qs, const double *ecoeff, int nnpg,
                  const double *shl, double omega, double xwidth, double zwidth, double rcoeff,
                  double dmax_in, double rkappa, double fcent, const double *xd, const double *zd,
                  double *det, double *vp, double *vs, double *qpp, double *qss, double *shg,
                  double complex *dampx, double complex *dampz)
{
    const double eps = 2.22e-16;
    for (int ix = 0; ix < nint; ix++) {
        for (int iz = 0; iz < nint; iz++) {
            int q = ix + nint * iz;
            double sf[4], dsf[8];
            if (orc_csf2d(xigll[ix], xigll[iz], sf, dsf)) return 1;
            double v_p = 0, v_s = 0, q_p = 0, q_s = 0, xdep = 0, zdep = 0;
            double dxdxi = 0, dxdeta = 0, dzdxi = 0, dzdeta = 0;
            for (int ia = 0; ia < 4; ia++) {
                int il = ieng_e[ia] - 1;
                double rlam = ecoeff[il], rmu = ecoeff[il + nnpg];
                double vpp = sqrt((rlam + 2.0 * rmu) / dens[il]);
                double vss = sqrt(rmu / dens[il]);
                v_p += sf[ia] * vpp;
                v_s += sf[ia] * vss;
                q_p += sf[ia] * qp[il];
                q_s += sf[ia] * qs[il];
                if (xd) { xdep += sf[ia] * xd[il]; zdep += sf[ia] * zd[il]; }
                dxdxi += dsf[2 * ia] * xlocs[il];
                dxdeta += dsf[2 * ia + 1] * xlocs[il];
                dzdxi += dsf[2 * ia] * zlocs[il];
                dzdeta += dsf[2 * ia + 1] * zlocs[il];
            }
            double rjac = dxdxi * dzdeta - dxdeta * dzdxi;
            if (fabs(rjac) <= eps || rjac < 0.0) return 1;
            det[q] = rjac; vp[q] = v_p; vs[q] = v_s; qpp[q] = q_p; qss[q] = q_s;
            if (dampx) {
                dampx[q] = xdep > eps ? orc_zdampfn(xdep, v_p, xwidth, dmax_in, rcoeff, rkappa, fcent, omega) : 1.0;
                dampz[q] = zdep > eps ? orc_zdampfn(zdep, v_p, zwidth, dmax_in, rcoeff, rkappa, fcent, omega) : 1.0;
            }
            for (int a = 0; a < nen; a++) {
                const double *s = shl + 3 * (a + nen * q);
                double *g = shg + 3 * (a + nen * q);
                g[2] = s[2];
                g[0] = (s[0] * dzdeta - s[1] * dzdxi) / rjac;
                g[1] = -(s[0] * dxdeta - s[1] * dxdxi) / rjac;
            }
        }
    }
    return 0;
}

/* estiff25.f:756-961 CMDAMP; with dampx = dampz = NULL it is KMDAMP
 * (estiff25.f:511-752), whose real/imaginary split is the same arithmetic
 * with unit stretching.  Density-normalised: per quadrature point
 * ZVP2 = VCMPLX(Vp)^2 W, ZVS2 = VCMPLX(Vs)^2 W (W = weights x DET), mass
 * coefficients -w^2 W + w^2 py^2 ZVS2 (x, z) and -w^2 W + w^2 py^2 ZVP2 (y),
 * real part into ME and imaginary part onto the diagonal of CE. */
void orc_kmdamp(int nen, int nint, double freq, double freq0, double py, const double *wts,
                const double *vp, const double *vs, const double *qpp, const double *qss,
                const double *det, const double *shg, const double complex *dampx,
                const double complex *dampz, double *me, double complex *ce)
{
    const double twopi = 6.2831853071795862;
    int nee = NDIM * nen;
    for (int i = 0; i < nee * nee; i++) { me[i] = 0.0; ce[i] = 0.0; }
    double omega = twopi * freq, wpy = omega * py, py2 = py * py, w2 = omega * omega;
    double complex cwpy = I * wpy;
    for (int ix = 0; ix < nint; ix++) {
        for (int iz = 0; iz < nint; iz++) {
            int q = ix + nint * iz;
            double weight = wts[ix] * wts[iz] * det[q];
            double complex ex = dampx ? dampx[q] : 1.0, ez = dampz ? dampz[q] : 1.0;
            double complex zvp = orc_vcmplx(freq, freq0, vp[q], qpp[q]);
            double complex zvs = orc_vcmplx(freq, freq0, vs[q], qss[q]);
            double complex zvp2 = zvp * zvp * weight, zvs2 = zvs * zvs * weight;
            double rm1 = -w2 * weight + w2 * py2 * creal(zvs2), qm1 = w2 * py2 * cimag(zvs2);
            double rm2 = -w2 * weight + w2 * py2 * creal(zvp2), qm2 = w2 * py2 * cimag(zvp2);
            const double *sq = shg + 3 * nen * q;
            for (int ib = 0; ib < nen; ib++) {
                double phib = sq[3 * ib + 2];
                double complex phibx = sq[3 * ib] * ex, phibz = sq[3 * ib + 1] * ez;
                double complex e11t1 = zvp2 * phibx, e11t2 = zvs2 * phibz;
                double complex e13t1 = zvp2 * phibz, e13t2 = zvs2 * phibx, e13t3 = -2.0 * zvs2 * phibz;
                double complex e22t1 = zvs2 * phibx, e22t2 = zvs2 * phibz;
                double complex e31t1 = zvp2 * phibx, e31t2 = zvs2 * phibz, e31t3 = -2.0 * zvs2 * phibx;
                double complex e33t1 = zvp2 * phibz, e33t2 = zvs2 * phibx;
                double complex e12t1 = cwpy * zvp2 * phib, e12t2 = -2.0 * cwpy * zvs2 * phib, e12t3 = -cwpy * zvs2 * phibx;
                double complex e21t1 = cwpy * zvp2 * phibx, e21t2 = -2.0 * cwpy * zvs2 * phibx, e21t3 = -cwpy * zvs2 * phib;
                double complex e23t1 = cwpy * zvp2 * phibz, e23t2 = -2.0 * cwpy * zvs2 * phibz, e23t3 = -cwpy * zvs2 * phib;
                double complex e32t1 = cwpy * zvp2 * phib, e32t2 = -2.0 * cwpy * zvs2 * phib, e32t3 = -cwpy * zvs2 * phibz;
                for (int ia = 0; ia < nen; ia++) {
                    double phia = sq[3 * ia + 2];
                    double complex phiax = sq[3 * ia] * ex, phiaz = sq[3 * ia + 1] * ez;
                    int ip = NDIM * ia, iq = NDIM * ib;
                    double rme11 = rm1 * phib * phia, qme11 = qm1 * phib * phia;
                    double rme22 = rm2 * phib * phia, qme22 = qm2 * phib * phia;
                    me[ip + nee * iq] += rme11;
                    me[ip + 1 + nee * (iq + 1)] += rme22;
                    me[ip + 2 + nee * (iq + 2)] += rme11;
                    double complex ce11 = e11t1 * phiax + e11t2 * phiaz;
                    double complex ce13 = e13t1 * phiax + e13t2 * phiaz + e13t3 * phiax;
                    double complex ce22 = e22t1 * phiax + e22t2 * phiaz;
                    double complex ce31 = e31t1 * phiaz + e31t2 * phiax + e31t3 * phiaz;
                    double complex ce33 = e33t1 * phiaz + e33t2 * phiax;
                    double complex ce12 = e12t1 * phiax + e12t2 * phiax + e12t3 * phia;
                    double complex ce21 = e21t1 * phia + e21t2 * phia + e21t3 * phiax;
                    double complex ce23 = e23t1 * phia + e23t2 * phia + e23t3 * phiaz;
                    double complex ce32 = e32t1 * phiaz + e32t2 * phiaz + e32t3 * phia;
                    ce[ip + nee * iq] += ce11 + I * qme11;
                    ce[ip + nee * (iq + 1)] -= ce12;
                    ce[ip + nee * (iq + 2)] += ce13;
                    ce[ip + 1 + nee * iq] += ce21;
                    ce[ip + 1 + nee * (iq + 1)] += ce22 + I * qme22;
                    ce[ip + 1 + nee * (iq + 2)] += ce23;
                    ce[ip + 2 + nee * iq] += ce31;
                    ce[ip + 2 + nee * (iq + 1)] -= ce32;
                    ce[ip + 2 + nee * (iq + 2)] += ce33 + I * qme11;
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* graph25.f:509-698 GENIEN.  Side nodes are deduplicated by a linear scan
 * (graph25.f:566-575), heap-sorted by x, z-sorted within groups of x that
 * agree to TOL (graph25.f:582-617), and numbered 1..NNPS in that order;
 * interior nodes follow in element order (graph25.f:672-675). */
typedef struct { double x, z; } xz_t;
static int cmp_x(const void *a, const void *b)
{
    double x = ((const xz_t *)a)->x, y = ((const xz_t *)b)->x;
    return (x > y) - (x < y);
}
static int cmp_z(const void *a, const void *b)
{
    double x = ((const xz_t *)a)->z, y = ((const xz_t *)b)->z;
    return (x > y) - (x < y);
}

static void node_xy(const orc_mesh *m, int e, double xi, double eta, double *x, double *z)
{
    double sf[4];
    csf2dnd(xi, eta, sf);
    double xx = 0, zz = 0;
    for (int k = 0; k < 4; k++) {
        int g = m->ieng[4 * e + k] - 1;
        xx += sf[k] * m->xlocs[g];
        zz += sf[k] * m->zlocs[g];
    }
    *x = xx; *z = zz;
}

/* sort.f:1046-1104 IBSECT8 with JOB=1, 1-based vec */
static int ibsect8(int n, double tol, double x, const double *vec1)
{
    const double *v = vec1 - 1;
    int r = -1, jl, ju, jm = 0;
    if (fabs(x - v[1]) < tol) return 1;
    if (fabs(x - v[n]) < tol) { r = n; goto back; }
    jl = 0; ju = n + 1;
    while (ju - jl > 1) {
        jm = (ju + jl) / 2;
        if ((v[n] >= v[1]) == (x >= v[jm])) jl = jm; else ju = jm;
    }
    if (jl <= 0) return -1;
    if (fabs(v[jl] - x) < tol) r = jl;
    else if (fabs(v[jm] - x) < tol) r = jm;
    else if (ju <= n && fabs(v[ju] - x) < tol) r = ju;
    else r = -1;
back:
    for (int i = r; i >= 2; i--)
        if (fabs(v[i] - v[i - 1]) > tol) return i;
    return r;
}

int orc_genien(const orc_mesh *m, const double *xipts, int *ien)
{
    const double tol = 1e-8;
    int nl = m->nord + 1, nen = nl * nl;
    int mnodes = m->nelem * (2 * (nl - 1) + 2 * nl);
    xz_t *nodes = (xz_t *)malloc(sizeof(xz_t) * (mnodes + 1));
    int inodes = 0;
    for (int e = 0; e < m->nelem; e++)
        for (int ie = 0; ie < nl; ie++)
            for (int ix = 0; ix < nl; ix++) {
                if (!(ix == 0 || ix == nl - 1 || ie == 0 || ie == nl - 1)) continue;
                double x, z;
                node_xy(m, e, xipts[ix], xipts[ie], &x, &z);
                int found = 0;
                for (int j = 0; j < inodes; j++)
                    if (fabs(nodes[j].x - x) < tol && fabs(nodes[j].z - z) < tol) { found = 1; break; }
                if (!found) { nodes[inodes].x = x; nodes[inodes].z = z; inodes++; }
            }
    int nn = inodes;
    qsort(nodes, nn, sizeof(xz_t), cmp_x);
    /* graph25.f:590-617, transcribed with 1-based indices */
    int jbeg = 1;
    for (int i = 1; i <= nn; i++) {
        if (i + 1 == nn) {
            if (jbeg != nn) qsort(nodes + jbeg - 1, nn - jbeg + 1, sizeof(xz_t), cmp_z);
            break;
        }
        if (i + 1 <= nn && fabs(nodes[i].x - nodes[i - 1].x) > tol) {
            if (i - jbeg + 1 > 1) qsort(nodes + jbeg - 1, i - jbeg + 1, sizeof(xz_t), cmp_z);
            jbeg = i + 1;
        }
    }
    int nnps = nn;
    double *xs = (double *)malloc(sizeof(double) * (nnps + 1));
    for (int i = 0; i < nnps; i++) xs[i] = nodes[i].x;
    xs[nnps] = HUGE_VAL;
    int iint = nnps;
    for (int e = 0; e < m->nelem; e++)
        for (int ie = 0; ie < nl; ie++)
            for (int ix = 0; ix < nl; ix++) {
                int a = ie * nl + ix;
                if (ix == 0 || ix == nl - 1 || ie == 0 || ie == nl - 1) {
                    double x, z;
                    node_xy(m, e, xipts[ix], xipts[ie], &x, &z);
                    int j0 = ibsect8(nnps + 1, tol, x, xs);
                    if (j0 < 0) { free(nodes); free(xs); return -3; }
                    int num = -1;
                    for (int j = j0; j <= nnps; j++)
                        if (fabs(nodes[j - 1].x - x) < tol && fabs(nodes[j - 1].z - z) < tol) { num = j; break; }
                    if (num < 0) { free(nodes); free(xs); return -4; }
                    ien[a + nen * e] = num;
                } else {
                    ien[a + nen * e] = ++iint;
                }
            }
    free(nodes); free(xs);
    return iint;
}

/* graph25.f:315-505 GENID */
int orc_genid(const orc_mesh *m, const double *xipts, const int *ien, int nnp, int *id)
{
    const double tol = 1e-10;
    const int lnull = -5;
    int nl = m->nord + 1, nen = nl * nl;
    for (int i = 0; i < NDIM * nnp; i++) id[i] = lnull;
#define FIXED(g) (m->cnnpg[2 * (g)] == 'F' && (m->cnnpg[2 * (g) + 1] == 'L' || m->cnnpg[2 * (g) + 1] == 'B' || m->cnnpg[2 * (g) + 1] == 'R'))
    for (int e = 0; e < m->nelem; e++) {
        if (m->cdomain[e] != 'A') continue;
        int icheck = 0, iav[3] = {0, 0, 0};
        for (int k = 0; k < 4; k++) {
            int iloc = m->ieng[4 * e + k];
            if (FIXED(iloc - 1)) {
                icheck++;
                iav[icheck - 1] = iloc;
                for (int j = 0; j < 4; j++) {
                    int jloc = m->ieng[4 * e + j];
                    if (FIXED(jloc - 1) && iloc != jloc) {
                        icheck++;
                        iav[icheck - 1] = jloc;
                        if (icheck == 3) break;
                    }
                }
                break;
            }
        }
        if (icheck == 0) continue;
        int ia1 = iav[0], ia2;
        double v1[3] = {0, 0, 0};
        for (int loop = 1; loop <= 2; loop++) {
            if (icheck > 1) {
                ia1 = iav[loop - 1];
                ia2 = (loop < 3) ? iav[loop] : 0;
                if (ia2 == 0) break;
                v1[0] = m->xlocs[ia2 - 1] - m->xlocs[ia1 - 1];
                v1[2] = m->zlocs[ia2 - 1] - m->zlocs[ia1 - 1];
            } else if (loop == 2) break;
            for (int ie = 0; ie < nl; ie++)
                for (int ix = 0; ix < nl; ix++) {
                    if (!(ix == 0 || ix == nl - 1 || ie == 0 || ie == nl - 1)) continue;
                    int inp = ien[ie * nl + ix + nen * e] - 1;
                    double x, z;
                    node_xy(m, e, xipts[ix], xipts[ie], &x, &z);
                    if (icheck == 1) {
                        if (fabs(x - m->xlocs[ia1 - 1]) < tol && fabs(z - m->zlocs[ia1 - 1]) < tol)
                            for (int i = 0; i < NDIM; i++) id[i + NDIM * inp] = 0;
                    } else {
                        double v2x = x - m->xlocs[ia1 - 1], v2z = z - m->zlocs[ia1 - 1];
                        double det = v1[0] * v2z - v1[2] * v2x;
                        if (fabs(det) < tol)
                            for (int i = 0; i < NDIM; i++) id[i + NDIM * inp] = 0;
                    }
                }
        }
    }
#undef FIXED
    int idof = 0;
    for (int e = 0; e < m->nelem; e++)
        for (int a = 0; a < nen; a++) {
            int inp = ien[a + nen * e] - 1;
            for (int i = 0; i < NDIM; i++)
                if (id[i + NDIM * inp] == lnull) id[i + NDIM * inp] = ++idof;
        }
    for (int i = 0; i < NDIM * nnp; i++)
        if (id[i] == lnull) return -1;
    return idof;
}

void orc_genlm(int nen, int nelem, const int *id, const int *ien, int *lm)
{
    for (int e = 0; e < nelem; e++)
        for (int a = 0; a < nen; a++)
            for (int i = 0; i < NDIM; i++)
                lm[i + NDIM * (a + nen * e)] = id[i + NDIM * (ien[a + nen * e] - 1)];
}

int orc_icmcon(int ngnod, int nnp, int nelem, const int *ien)
{
    int *cnt = (int *)calloc(nnp, sizeof(int));
    for (int e = 0; e < nelem; e++)
        for (int k = 0; k < ngnod; k++) cnt[ien[k + ngnod * e] - 1]++;
    int mx = 0;
    for (int i = 0; i < nnp; i++) if (cnt[i] > mx) mx = cnt[i];
    free(cnt);
    return mx;
}

void orc_connect(int ngnod, int nnp, int nelem, int ncon, const int *ien, int *mconn)
{
    memset(mconn, 0, sizeof(int) * (size_t)nnp * ncon);
    for (int e = 0; e < nelem; e++)
        for (int k = 0; k < ngnod; k++) {
            int inp = ien[k + ngnod * e] - 1;
            for (int c = 0; c < ncon; c++)
                if (mconn[inp + (size_t)nnp * c] == 0) { mconn[inp + (size_t)nnp * c] = e + 1; break; }
        }
}

static int cmp_int(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* graph25.f:136-266 GENCRS_QUICK (rows built in first-encounter order, then
 * permuted to DOF order) */
long orc_gencrs(int nen, int nnp, int nelem, int ndof, int ncon, const int *mconn,
                const int *ien, const int *lm, int *irptr, int *jcptr, long cap)
{
    int nwork = NDIM * ncon * nen + 1;
    int *work = (int *)malloc(sizeof(int) * (nwork + 1));
    int *perm = (int *)calloc(ndof + 1, sizeof(int));
    long *rw = (long *)malloc(sizeof(long) * (ndof + 2));
    int *jw = (int *)malloc(sizeof(int) * cap);
    long izero = 0;
    int ifound = 0;
    rw[1] = 1;
    for (int e = 0; e < nelem; e++)
        for (int a = 0; a < nen; a++) {
            int inp = ien[a + nen * e] - 1;
            for (int i = 0; i < NDIM; i++) {
                int idof = lm[i + NDIM * (a + nen * e)];
                if (idof <= 0 || perm[idof]) continue;
                int nrr = 0;
                for (int c = 0; c < ncon; c++) {
                    int je = mconn[inp + (size_t)nnp * c];
                    if (je == 0) break;
                    for (int b = 0; b < nen; b++)
                        for (int j = 0; j < NDIM; j++) {
                            int jd = lm[j + NDIM * (b + nen * (je - 1))];
                            if (jd > 0) work[nrr++] = jd;
                        }
                }
                qsort(work, nrr, sizeof(int), cmp_int);
                work[nrr] = 0;
                for (int r = 0; r < nrr; r++)
                    if (work[r] != work[r + 1]) {
                        if (izero >= cap) { free(work); free(perm); free(rw); free(jw); return -2; }
                        jw[izero++] = work[r];
                    }
                ifound++;
                perm[idof] = ifound;
                rw[ifound + 1] = izero + 1;
            }
        }
    long iz = 0;
    irptr[0] = 1;
    for (int idof = 1; idof <= ndof; idof++) {
        int irow = perm[idof];
        for (long j = rw[irow]; j < rw[irow + 1]; j++) jcptr[iz++] = jw[j - 1];
        irptr[idof] = (int)(iz + 1);
    }
    free(work); free(perm); free(rw); free(jw);
    return iz;
}

/* bielak_utils.f:934-1025 CNMESH */
int orc_cnmesh(int nen, int ndof, int nelem, const char *cdomain, const int *lm, char *cnp)
{
    int ierr = 0;
    memset(cnp, ' ', ndof);
    for (int e = 0; e < nelem; e++) if (cdomain[e] == 'A')
        for (int a = 0; a < nen * NDIM; a++) { int d = lm[a + NDIM * nen * e]; if (d > 0) cnp[d - 1] = 'A'; }
    for (int e = 0; e < nelem; e++) if (cdomain[e] == 'E')
        for (int a = 0; a < nen * NDIM; a++) { int d = lm[a + NDIM * nen * e]; if (d > 0 && cnp[d - 1] == ' ') cnp[d - 1] = 'E'; }
    for (int e = 0; e < nelem; e++) if (cdomain[e] == 'I')
        for (int a = 0; a < nen * NDIM; a++) {
            int d = lm[a + NDIM * nen * e];
            if (d <= 0) continue;
            if (cnp[d - 1] == ' ') cnp[d - 1] = 'I';
            else if (cnp[d - 1] == 'E') cnp[d - 1] = 'B';
            else if (cnp[d - 1] == 'A') ierr = 1;
        }
    for (int d = 0; d < ndof; d++) if (cnp[d] == ' ') ierr = 2;
    return ierr;
}

/* ------------------------------------------------------------------------ */
/* sort.f:1006-1040 IBSECT on a sorted, 1-based column list (returns 1-based) */
static int ibsect(int n, int ix, const int *v)
{
    int lo = 0, hi = n - 1;
    while (lo <= hi) {
        int mid = (lo + hi) >> 1;
        if (v[mid] == ix) return mid + 1;
        if (v[mid] < ix) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

typedef struct {
    int nl, nen, nee, nint;
    double *xipts, *xigll, *wgll, *shl, *shg, *det, *rho, *dmat, *me, *vq;
    double complex *ke, *dampx, *dampz;
} elem_ws;

static void ws_init(elem_ws *w, int nord)
{
    w->nl = nord + 1; w->nen = w->nl * w->nl; w->nee = NDIM * w->nen;
    w->nint = 2 * nord + 5;   /* asmble.f90:110-112 */
    int q = w->nint * w->nint;
    w->xipts = (double *)malloc(sizeof(double) * w->nl);
    w->xigll = (double *)malloc(sizeof(double) * w->nint);
    w->wgll = (double *)malloc(sizeof(double) * w->nint);
    w->shl = (double *)malloc(sizeof(double) * 3 * w->nen * q);
    w->shg = (double *)malloc(sizeof(double) * 3 * w->nen * q);
    w->det = (double *)malloc(sizeof(double) * q);
    w->rho = (double *)malloc(sizeof(double) * q);
    w->dmat = (double *)malloc(sizeof(double) * 2 * q);
    w->vq = (double *)malloc(sizeof(double) * q);
    w->me = (double *)malloc(sizeof(double) * w->nee * w->nee);
    w->ke = (double complex *)malloc(sizeof(double complex) * w->nee * w->nee);
    w->dampx = (double complex *)malloc(sizeof(double complex) * q);
    w->dampz = (double complex *)malloc(sizeof(double complex) * q);
    orc_dshl(w->nl, w->nint, w->xipts, w->xigll, w->wgll, w->shl);
}

static void ws_free(elem_ws *w)
{
    free(w->xipts); free(w->xigll); free(w->wgll); free(w->shl); free(w->shg);
    free(w->det); free(w->rho); free(w->dmat); free(w->vq); free(w->me); free(w->ke);
    free(w->dampx); free(w->dampz);
}

/* asmble.f90:140-212: element dispatch, 'A' -> CJACCQ+CM25ISOQ, else CJMK_TVI+KM25ISOQ */
static int element_matrix(const orc_asm *a, elem_ws *w, int e, double omega, double py)
{
    const int *ie = a->ieng + 4 * e;
    int q = w->nint * w->nint;
    if (a->ldisp) {   /* asmble.f90:159-168 and 187-201 */
        const double freq = omega / 6.2831853071795862;
        const int pml = a->cdomain[e] == 'A';
        if (orc_cjkm_qual(w->nen, w->nint, ie, w->xigll, a->xlocs, a->zlocs, a->dens, a->qp, a->qs,
                          a->ecoeff, a->nnpg, w->shl, omega, a->xwidth, a->zwidth, a->rcoeff, a->dmax,
                          a->kappa, a->fcent, pml ? a->xd : NULL, pml ? a->zd : NULL, w->det,
                          w->dmat, w->dmat + q, w->rho, w->vq, w->shg, pml ? w->dampx : NULL,
                          pml ? w->dampz : NULL)) return 1;
        orc_kmdamp(w->nen, w->nint, freq, a->freq0, py, w->wgll, w->dmat, w->dmat + q, w->rho, w->vq,
                   w->det, w->shg, pml ? w->dampx : NULL, pml ? w->dampz : NULL, w->me, w->ke);
        return 0;
    }
    if (a->cdomain[e] != 'A') {
        if (orc_cjmk_tvi(w->nen, w->nint, ie, w->xigll, a->xlocs, a->zlocs, a->dens, a->ecoeff,
                         a->nnpg, w->shl, w->det, w->rho, w->dmat, w->shg)) return 1;
        orc_km25isoq(w->nen, w->nint, omega, py, w->wgll, w->dmat, w->dmat + q, w->rho, w->det,
                     w->shg, w->me, w->ke);
    } else {
        if (orc_cjaccq(w->nen, w->nint, omega, a->xwidth, a->zwidth, a->rcoeff, a->dmax, a->kappa,
                       a->fcent, ie, w->xigll, a->xd, a->zd, a->xlocs, a->zlocs, a->dens, a->ecoeff,
                       a->nnpg, w->shl, w->det, w->rho, w->dmat, w->shg, w->dampx, w->dampz)) return 1;
        orc_cm25isoq(w->nen, w->nint, omega, py, w->wgll, w->dmat, w->dmat + q, w->rho, w->det,
                     w->shg, w->dampx, w->dampz, w->me, w->ke);
    }
    return 0;
}

static int asmble_common(const orc_asm *a, double freq, double py, float complex *cdist,
                         double complex *sz)
{
    const double twopi = 6.2831853071795862;
    double omega = twopi * freq;
    elem_ws w;
    ws_init(&w, a->nord);
    long nz = a->irptr[a->ndof] - 1;
    if (cdist) for (long i = 0; i < nz; i++) cdist[i] = 0.0f;
    if (sz) for (long i = 0; i < nz; i++) sz[i] = 0.0;
    int ierr = 0;
    for (int e = 0; e < a->nelem && !ierr; e++) {
        if (element_matrix(a, &w, e, omega, py)) { ierr = 1; break; }
        const int *lm = a->lm + NDIM * w.nen * e;
        /* asmble.f90:279-312 ASMBLE_CDIST */
        for (int ip = 0; ip < w.nee; ip++) {
            int idof = lm[ip];
            if (idof == 0) continue;
            int jbeg = a->irptr[idof - 1], jend = a->irptr[idof] - 1;
            int nvar = jend - jbeg + 1;
            for (int iq = 0; iq < w.nee; iq++) {
                int jdof = lm[iq];
                if (jdof == 0) continue;
                int indx = ibsect(nvar, jdof, a->jcptr + jbeg - 1);
                if (indx <= 0) { ierr = 1; break; }
                long iz = jbeg - 1 + indx - 1;
                double r1 = w.me[ip + w.nee * iq] + creal(w.ke[ip + w.nee * iq]);
                double q1 = cimag(w.ke[ip + w.nee * iq]);
                if (cdist) {
                    float cr = (float)r1, ci = (float)q1;
                    float ar = crealf(cdist[iz]) + cr, ai = cimagf(cdist[iz]) + ci;
                    cdist[iz] = ar + I * ai;
                }
                if (sz) sz[iz] += r1 + I * q1;
            }
        }
    }
    ws_free(&w);
    return ierr;
}

int orc_asmble_cdist(const orc_asm *a, double freq, double py, float complex *cdist)
{
    return asmble_common(a, freq, py, cdist, NULL);
}

int orc_asmble_z(const orc_asm *a, double freq, double py, double complex *s)
{
    return asmble_common(a, freq, py, NULL, s);
}

void orc_spmv_z(int ndof, const int *irptr, const int *jcptr, const double complex *s,
                const double complex *x, double complex *y)
{
    for (int i = 0; i < ndof; i++) {
        double complex acc = 0.0;
        for (int j = irptr[i] - 1; j < irptr[i + 1] - 1; j++) acc += s[j] * x[jcptr[j] - 1];
        y[i] = acc;
    }
}

void orc_spmv_c(int ndof, const int *irptr, const int *jcptr, const float complex *s,
                const float complex *x, float complex *y)
{
    for (int i = 0; i < ndof; i++) {
        float ar = 0.f, ai = 0.f;
        for (int j = irptr[i] - 1; j < irptr[i + 1] - 1; j++) {
            float sr = crealf(s[j]), si = cimagf(s[j]);
            float xr = crealf(x[jcptr[j] - 1]), xi = cimagf(x[jcptr[j] - 1]);
            ar += sr * xr - si * xi;
            ai += sr * xi + si * xr;
        }
        y[i] = ar + I * ai;
    }
}

/* asmble.f90:318-377 ASMB_PEFF without the reduce (one rank owns every row) */
void orc_asmb_peff(int ndof, const char *cnp, const int *irptr, const int *jcptr,
                   const float complex *cdist, const float complex *ue, float complex *peff)
{
    for (int i = 0; i < ndof; i++) peff[i] = 0.0f;
    for (int i = 0; i < ndof; i++) {
        char c1 = cnp[i];
        if (c1 != 'B' && c1 != 'E') continue;
        float br = crealf(peff[i]), bi = cimagf(peff[i]);
        for (int j = irptr[i] - 1; j < irptr[i + 1] - 1; j++) {
            int jd = jcptr[j] - 1;
            char c2 = cnp[jd];
            float sr = crealf(cdist[j]), si = cimagf(cdist[j]);
            float ur = crealf(ue[jd]), ui = cimagf(ue[jd]);
            float pr = sr * ur - si * ui, pi = sr * ui + si * ur;
            if (c1 == 'B' && c2 == 'E') { br -= pr; bi -= pi; }
            else if (c1 == 'E' && c2 == 'B') { br += pr; bi += pi; }
        }
        peff[i] = br + I * bi;
    }
}

/* ------------------------------------------------------------------------ */
/* djac.f:1-139 DJACAB */
int orc_djacab(int nen, int nint, int inpg, const int *ieng_e, const double *xigll,
               const double *xlocs, const double *zlocs, const double *dens,
               const double *ecoeff, int nnpg, const double *shl,
               double *det, double *dab, double *shg)
{
    for (int ix = 0; ix < nint; ix++)
        for (int iz = 0; iz < nint; iz++) {
            double sf[4], rjac, a, b, c, d;
            int err = jacobian_at(xigll[ix], xigll[iz], ieng_e, xlocs, zlocs, sf, &rjac, &a, &b, &c, &d);
            int q = ix + nint * iz;
            double d1 = 0, d2 = 0;
            for (int k = 0; k < 4; k++) {
                int g = ieng_e[k];
                double al = 0, be = 0;
                if (g == inpg) {
                    al = sqrt((ecoeff[g - 1] + 2.0 * ecoeff[g - 1 + nnpg]) / dens[g - 1]);
                    be = sqrt(ecoeff[g - 1 + nnpg] / dens[g - 1]);
                }
                d1 += 2.0 * al * sf[k];
                d2 += 2.0 * be * sf[k];
            }
            dab[q] = d1; dab[q + nint * nint] = d2;
            det[q] = rjac;
            if (err) return 1;
            fill_shg(nen, shl + 3 * nen * q, shg + 3 * nen * q, rjac, a, b, c, d);
        }
    return 0;
}

/* dstiff25.f:498-662 DKM_VPVS */
void orc_dkm_vpvs(int nen, int nint, double omega, double py, const double *wts,
                  const double *dvp, const double *dvs, const double *det, const double *shg,
                  double *dme, double complex *dke)
{
    int nee = NDIM * nen;
    for (int i = 0; i < nee * nee; i++) { dme[i] = 0.0; dke[i] = 0.0; }
    double wpy = omega * py, py2 = py * py, w2 = omega * omega;
    for (int ix = 0; ix < nint; ix++)
        for (int iz = 0; iz < nint; iz++) {
            int q = ix + nint * iz;
            double weight = wts[ix] * wts[iz] * det[q];
            double drvp = dvp[q] * weight, drvs = dvs[q] * weight;
            double rm1 = w2 * py2 * drvs, rm2 = w2 * py2 * drvp;
            const double *sq = shg + 3 * nen * q;
            for (int ib = 0; ib < nen; ib++) {
                double phib = sq[3 * ib + 2], phibx = sq[3 * ib], phibz = sq[3 * ib + 1];
                double rm11t1 = rm1 * phib, rm22t1 = rm2 * phib;
                double rk11t1 = drvp * phibx, rk11t2 = drvs * phibz;
                double rk13t1 = drvp * phibz, rk13t2 = drvs * phibx, rk13t3 = -2.0 * drvs * phibz;
                double rk22t1 = drvs * phibx, rk22t2 = drvs * phibz;
                double rk31t1 = drvp * phibx, rk31t2 = drvs * phibz, rk31t3 = -2.0 * drvs * phibx;
                double rk33t1 = drvp * phibz, rk33t2 = drvs * phibx;
                double rk12t1 = wpy * drvp * phib, rk12t2 = -2.0 * wpy * drvs * phib, rk12t3 = -wpy * drvs * phibx;
                double rk21t1 = wpy * drvp * phibx, rk21t2 = -2.0 * wpy * drvs * phibx, rk21t3 = -wpy * drvs * phib;
                double rk23t1 = wpy * drvp * phibz, rk23t2 = -2.0 * wpy * drvs * phibz, rk23t3 = -wpy * drvs * phib;
                double rk32t1 = wpy * drvp * phib, rk32t2 = -2.0 * wpy * drvs * phib, rk32t3 = -wpy * drvs * phibz;
                for (int ia = 0; ia < nen; ia++) {
                    double phia = sq[3 * ia + 2], phiax = sq[3 * ia], phiaz = sq[3 * ia + 1];
                    double rme11 = rm11t1 * phia, rme22 = rm22t1 * phia;
                    double rke11 = rk11t1 * phiax + rk11t2 * phiaz;
                    double rke13 = rk13t1 * phiax + rk13t2 * phiaz + rk13t3 * phiax;
                    double rke22 = rk22t1 * phiax + rk22t2 * phiaz;
                    double rke31 = rk31t1 * phiaz + rk31t2 * phiax + rk31t3 * phiaz;
                    double rke33 = rk33t1 * phiaz + rk33t2 * phiax;
                    double rke12 = rk12t1 * phiax + rk12t2 * phiax + rk12t3 * phia;
                    double rke21 = rk21t1 * phia + rk21t2 * phia + rk21t3 * phiax;
                    double rke23 = rk23t1 * phia + rk23t2 * phia + rk23t3 * phiaz;
                    double rke32 = rk32t1 * phiaz + rk32t2 * phiaz + rk32t3 * phia;
                    int ip = NDIM * ia, iq = NDIM * ib;
                    dme[ip + nee * iq] += rme11;
                    dme[ip + 1 + nee * (iq + 1)] += rme22;
                    dme[ip + 2 + nee * (iq + 2)] += rme11;
                    dke[ip + nee * iq] += rke11;
                    dke[ip + 1 + nee * (iq + 1)] += rke22;
                    dke[ip + 2 + nee * (iq + 2)] += rke33;
                    dke[ip + nee * (iq + 2)] += rke13;
                    dke[ip + 2 + nee * iq] += rke31;
                    dke[ip + nee * (iq + 1)] -= rke12;
                    dke[ip + 1 + nee * iq] += rke21;
                    dke[ip + 1 + nee * (iq + 2)] += rke23;
                    dke[ip + 2 + nee * (iq + 1)] -= rke32;
                }
            }
        }
}

/* djac.f:143-288 DJACAB2: DJACAB's Jacobian, SHG and DAB plus the collocated
 * second-derivative field DAB2(q,1) = DAB2(q,2) = 2 N_INPG(q). */
int orc_djacab2(int nen, int nint, int inpg, const int *ieng_e, const double *xigll,
                const double *xlocs, const double *zlocs, const double *dens,
                const double *ecoeff, int nnpg, const double *shl,
                double *det, double *dab, double *dab2, double *shg)
{
    int nq = nint * nint;
    for (int ix = 0; ix < nint; ix++)
        for (int iz = 0; iz < nint; iz++) {
            double sf[4], rjac, a, b, c, d;
            int err = jacobian_at(xigll[ix], xigll[iz], ieng_e, xlocs, zlocs, sf, &rjac, &a, &b, &c, &d);
            int q = ix + nint * iz;
            double d1 = 0, d2 = 0, e1 = 0, e2 = 0;
            for (int k = 0; k < 4; k++) {
                int g = ieng_e[k];
                double al = 0, be = 0, da2 = 0, db2 = 0;
                if (g == inpg) {
                    da2 = 2.0; db2 = 2.0;
                    al = sqrt((ecoeff[g - 1] + 2.0 * ecoeff[g - 1 + nnpg]) / dens[g - 1]);
                    be = sqrt(ecoeff[g - 1 + nnpg] / dens[g - 1]);
                }
                d1 += 2.0 * al * sf[k];
                d2 += 2.0 * be * sf[k];
                e1 += da2 * sf[k];
                e2 += db2 * sf[k];
            }
            dab[q] = d1; dab[q + nq] = d2;
            dab2[q] = e1; dab2[q + nq] = e2;
            det[q] = rjac;
            if (err) return 1;
            fill_shg(nen, shl + 3 * nen * q, shg + 3 * nen * q, rjac, a, b, c, d);
        }
    return 0;
}

/* dstiff25.f:1-203 DABMKISOB25: derivatives of the density-normalised element
 * matrices with respect to alpha (Vp, lisp) and beta (Vs, liss) collocated on
 * one anchor through DAB; the ky coupling terms are imaginary.  Outputs are
 * NEE x NEE column-major; the explicit zero stores of the reference
 * (DMDA x/z diagonal, DMDB y diagonal, DMDB z = DMDB x) are kept. */
void orc_dabmkisob25(int nen, int nint, int lisp, int liss, double omega, double py,
                     const double *wts, const double *det, const double *dab, const double *shg,
                     double *dmda, double *dmdb, double complex *dkda, double complex *dkdb)
{
    int nee = NDIM * nen, nq = nint * nint;
    for (int i = 0; i < nee * nee; i++) { dmda[i] = 0.0; dmdb[i] = 0.0; dkda[i] = 0.0; dkdb[i] = 0.0; }
    double wpy = omega * py, wpy2 = (omega * py) * (omega * py);
    for (int ix = 0; ix < nint; ix++)
        for (int iz = 0; iz < nint; iz++) {
            int q = ix + nint * iz;
            double weight = wts[ix] * wts[iz] * det[q];
            double dalpha = dab[q] * weight, dbeta = dab[q + nq] * weight;
            double dmassda = wpy2 * dalpha, dmassdb = wpy2 * dbeta;
            const double *sq = shg + 3 * nen * q;
            for (int ib = 0; ib < nen; ib++) {
                double phib = sq[3 * ib + 2], phibx = sq[3 * ib], phibz = sq[3 * ib + 1];
                double dmdat1 = 0, dk11a = 0, dk13a = 0, dk31a = 0, dk33a = 0;
                double dk12a = 0, dk21a = 0, dk23a = 0, dk32a = 0;
                double dmdbt1 = 0, b11 = 0, b13a = 0, b13b = 0, b22a = 0, b22b = 0, b31a = 0, b31b = 0, b33 = 0;
                double b12a = 0, b12b = 0, b21a = 0, b21b = 0, b23a = 0, b23b = 0, b32a = 0, b32b = 0;
                if (lisp) {
                    dmdat1 = dmassda * phib;
                    dk11a = dalpha * phibx; dk13a = dalpha * phibz;
                    dk31a = dalpha * phibx; dk33a = dalpha * phibz;
                    dk12a = wpy * dalpha * phib; dk21a = wpy * dalpha * phibx;
                    dk23a = wpy * dalpha * phibz; dk32a = wpy * dalpha * phib;
                }
                if (liss) {
                    dmdbt1 = dmassdb * phib;
                    b11 = dbeta * phibz;
                    b13a = dbeta * phibx; b13b = -2.0 * dbeta * phibz;
                    b22a = dbeta * phibx; b22b = dbeta * phibz;
                    b31a = dbeta * phibz; b31b = -2.0 * dbeta * phibx;
                    b33 = dbeta * phibx;
                    b12a = 2.0 * wpy * dbeta * phib; b12b = wpy * dbeta * phibx;
                    b21a = 2.0 * wpy * dbeta * phibx; b21b = wpy * dbeta * phib;
                    b23a = 2.0 * wpy * dbeta * phibz; b23b = wpy * dbeta * phib;
                    b32a = 2.0 * wpy * dbeta * phib; b32b = wpy * dbeta * phibz;
                }
                for (int ia = 0; ia < nen; ia++) {
                    double phia = sq[3 * ia + 2], phiax = sq[3 * ia], phiaz = sq[3 * ia + 1];
                    int ip = NDIM * ia, iq = NDIM * ib;
#define AT(i, j) ((ip + (i)) + nee * (iq + (j)))
                    if (lisp) {
                        dmda[AT(0, 0)] = 0.0;
                        dmda[AT(1, 1)] = dmda[AT(1, 1)] + dmdat1 * phia;
                        dmda[AT(2, 2)] = 0.0;
                    }
                    if (liss) {
                        dmdb[AT(0, 0)] = dmdb[AT(0, 0)] + dmdbt1 * phia;
                        dmdb[AT(1, 1)] = 0.0;
                        dmdb[AT(2, 2)] = dmdb[AT(0, 0)];
                    }
                    if (lisp) {
                        dkda[AT(0, 0)] += phiax * dk11a;
                        dkda[AT(0, 1)] -= I * (phiax * dk12a);
                        dkda[AT(0, 2)] += phiax * dk13a;
                        dkda[AT(1, 0)] += I * (phia * dk21a);
                        dkda[AT(1, 2)] += I * (phia * dk23a);
                        dkda[AT(2, 0)] += phiaz * dk31a;
                        dkda[AT(2, 1)] -= I * (phiaz * dk32a);
                        dkda[AT(2, 2)] += phiaz * dk33a;
                    }
                    if (liss) {
                        dkdb[AT(0, 0)] += phiaz * b11;
                        dkdb[AT(0, 1)] += I * (phiax * b12a + phia * b12b);
                        dkdb[AT(0, 2)] += phiaz * b13a + phiax * b13b;
                        dkdb[AT(1, 0)] -= I * (phia * b21a + phiax * b21b);
                        dkdb[AT(1, 1)] += phiax * b22a + phiaz * b22b;
                        dkdb[AT(1, 2)] -= I * (phia * b23a + phiaz * b23b);
                        dkdb[AT(2, 0)] += phiax * b31a + phiaz * b31b;
                        dkdb[AT(2, 1)] += I * (phiaz * b32a + phia * b32b);
                        dkdb[AT(2, 2)] += phiax * b33;
                    }
#undef AT
                }
            }
        }
}

/* dstiff25.f:208-494 DABM2KISOB25: the first-derivative matrices of
 * DABMKISOB25 and the same expressions on DAB2 (second derivatives).  Each
 * output is computed by the statements DABMKISOB25 uses for it, so this is
 * DABMKISOB25 applied to DAB and to DAB2. */
void orc_dabm2kisob25(int nen, int nint, int lisp, int liss, double omega, double py,
                      const double *wts, const double *det, const double *dab, const double *dab2,
                      const double *shg, double *dmda, double *dmdb, double complex *dkda,
                      double complex *dkdb, double *dmda2, double *dmdb2, double complex *dkda2,
                      double complex *dkdb2)
{
    orc_dabmkisob25(nen, nint, lisp, liss, omega, py, wts, det, dab, shg, dmda, dmdb, dkda, dkdb);
    orc_dabmkisob25(nen, nint, lisp, liss, omega, py, wts, det, dab2, shg, dmda2, dmdb2, dkda2, dkdb2);
}

/* ------------------------------------------------------------------------ */
static int cnnpg_is(const char *cnnpg, int g, const char *s)
{
    return cnnpg[2 * g] == s[0] && cnnpg[2 * g + 1] == s[1];
}

/* gengmask.f90:1-262 GENGMASK (no .mask file: every weight is 1; one partition) */
int orc_gengmask(const orc_asm *a, const char *cnnpg, int nvinv, int linrec, int lkbdry,
                 int nrec, const int *mrdof1, orc_inv *inv)
{
    int nl = a->nord + 1, nen = nl * nl;
    memset(inv, 0, sizeof(*inv));
    inv->nvinv = nvinv;
    inv->ncon = orc_icmcon(4, a->nnpg, a->nelem, a->ieng);
    inv->mconn = (int *)malloc(sizeof(int) * (size_t)a->nnpg * inv->ncon);
    orc_connect(4, a->nnpg, a->nelem, inv->ncon, a->ieng, inv->mconn);
    inv->maskg = (int *)calloc(a->nnpg, sizeof(int));
    int inpinv = 0, lshift = 0;
    for (int g = 0; g < a->nnpg; g++) {
        int keep = 1, e = inv->mconn[g];
        if (e > 0 && a->cdomain[e - 1] != 'I') keep = 0;
        if (keep && (cnnpg_is(cnnpg, g, "II") || cnnpg_is(cnnpg, g, "BI") || cnnpg_is(cnnpg, g, "FS")))
            inv->maskg[g] = ++inpinv;
    }
    inv->nnpinv = inpinv;
    int ia2iae[4] = {1, nl, nl * nl, nl * (nl - 1) + 1};
    for (int e = 0; e < a->nelem; e++)
        for (int k = 0; k < 4; k++) {
            int g = a->ieng[4 * e + k] - 1;
            if (inv->maskg[g] <= 0) continue;
            int iae = ia2iae[k] - 1;
            int drop = 0;
            for (int i = 0; i < NDIM; i++)
                if (a->lm[i + NDIM * (iae + nen * e)] <= 0) drop = 1;
            if (!drop && !linrec)
                for (int r = 0; r < nrec; r++)
                    if (a->lm[NDIM * (iae + nen * e)] == mrdof1[r]) { drop = 1; break; }
            if (drop) { inv->nnpinv--; inv->maskg[g] = 0; lshift = 1; }
        }
    if (!lkbdry)
        for (int e = 0; e < a->nelem; e++) {
            int bdr = 0;
            for (int k = 0; k < 4; k++) {
                int g = a->ieng[4 * e + k] - 1;
                if (!cnnpg_is(cnnpg, g, "II") && !cnnpg_is(cnnpg, g, "FS")) bdr = 1;
            }
            if (bdr)
                for (int k = 0; k < 4; k++) {
                    int g = a->ieng[4 * e + k] - 1;
                    if (inv->maskg[g] > 0) { inv->maskg[g] = 0; inv->nnpinv--; lshift = 1; }
                }
        }
    if (inv->nnpinv < 1) return 1;
    if (lshift) {
        inpinv = 0;
        for (int g = 0; g < a->nnpg; g++) if (inv->maskg[g] > 0) inv->maskg[g] = ++inpinv;
    }
    inv->na35 = inv->nnpinv * nvinv;
    return 0;
}

/* gengmask.f90:962-1198 GRADPTRS (MYID = 0, one partition) */
int orc_gradptrs(const orc_asm *a, orc_inv *inv)
{
    int nl = a->nord + 1, nen = nl * nl, nee = NDIM * nen;
    inv->nnpgl = inv->nnpinv * inv->nvinv;
    long nwork = (long)a->nnpg * inv->ncon * nee;
    int *irow = (int *)malloc(sizeof(int) * (inv->ncon * nee + 1));
    inv->jcsc = (int *)malloc(sizeof(int) * (inv->nnpgl + 1));
    inv->icsc = (int *)malloc(sizeof(int) * nwork);
    inv->mygrad = (int *)malloc(sizeof(int) * inv->nnpgl);
    int inzfl = 0, inpgl = 0, combig = 0;
    inv->jcsc[0] = 1;
    for (int g = 0; g < a->nnpg; g++) {
        int inpinv = inv->maskg[g];
        if (inpinv <= 0) continue;
        int nn = 0;
        for (int c = 0; c < inv->ncon; c++) {
            int e = inv->mconn[g + (size_t)a->nnpg * c];
            if (e <= 0) break;
            for (int ae = 0; ae < nen; ae++)
                for (int i = 0; i < NDIM; i++) {
                    int d = a->lm[i + NDIM * (ae + nen * (e - 1))];
                    if (d > 0) irow[nn++] = d;
                }
        }
        qsort(irow, nn, sizeof(int), cmp_int);
        irow[nn] = 0;
        for (int r = 0; r < nn; r++)
            if (irow[r + 1] != irow[r]) {
                if (inzfl >= nwork) { free(irow); return 1; }   /* 'ICSC_FDIST too small' */
                inv->icsc[inzfl++] = irow[r];
            }
        inv->mygrad[inpgl] = (inpinv - 1) * inv->nvinv + 1;
        inv->jcsc[inpgl + 1] = inzfl + 1;
        int r1 = inv->jcsc[inpgl], r2 = inv->jcsc[inpgl + 1] - 1;
        if (r2 - r1 + 1 > combig) combig = r2 - r1 + 1;
        for (int iv = 2; iv <= inv->nvinv; iv++) {
            for (int r = r1; r <= r2; r++) {
                if (inzfl >= nwork) { free(irow); return 1; }
                inv->icsc[inzfl++] = inv->icsc[r - 1];
            }
            inpgl++;
            inv->mygrad[inpgl] = (inpinv - 1) * inv->nvinv + iv;
            inv->jcsc[inpgl + 1] = inzfl + 1;
        }
        inpgl++;
    }
    inv->nz_fdist = inzfl;
    inv->mbufrhs = combig;
    free(irow);
    return (inpgl == inv->nnpgl) ? 0 : 1;
}

/* gengmask.f90:758-841 ELEM_WTS */
int orc_elem_wts(const orc_asm *a, orc_inv *inv)
{
    inv->elem_wts = (double *)calloc(a->nelem, sizeof(double));
    char *init = (char *)calloc(a->nelem, 1);
    double wtmax = 0.0;
    for (int g = 0; g < a->nnpg; g++) {
        if (inv->maskg[g] <= 0) continue;
        for (int c = 0; c < inv->ncon; c++) {
            int e = inv->mconn[g + (size_t)a->nnpg * c];
            if (e <= 0) break;
            e--;
            if (init[e]) continue;
            double rloc = 0.0;
            for (int k = 0; k < 4; k++) {
                int j1 = a->ieng[4 * e + k] - 1, j2 = a->ieng[4 * e + (k + 1) % 4] - 1;
                rloc += a->xlocs[j1] * a->zlocs[j2] - a->xlocs[j2] * a->zlocs[j1];
            }
            if (rloc <= 0.0) rloc = -rloc;
            rloc *= 0.5;
            inv->elem_wts[e] = rloc;
            init[e] = 1;
            if (rloc > wtmax) wtmax = rloc;
        }
    }
    for (int e = 0; e < a->nelem; e++)
        if (inv->elem_wts[e] > 0.0) inv->elem_wts[e] = wtmax / inv->elem_wts[e];
    free(init);
    return wtmax > 0.0 ? 0 : 1;
}

void orc_inv_free(orc_inv *inv)
{
    free(inv->mconn); free(inv->maskg); free(inv->mygrad); free(inv->jcsc); free(inv->icsc);
    free(inv->elem_wts);
    memset(inv, 0, sizeof(*inv));
}

static inline float complex cmulf_ref(float complex a, float complex b)
{
    float ar = crealf(a), ai = cimagf(a), br = crealf(b), bi = cimagf(b);
    return (ar * br - ai * bi) + I * (ar * bi + ai * br);
}

/* hessloc.f90:1-133 CFMAT_DIST driving vforce25.f:1-369 VFORCE25 (NVINV = 1).
 * VFORCE25 clears FORCEV inside its loop over the node's elements
 * (vforce25.f:252-256 sit inside DO 200), so each column holds the product of
 * the last connected element only; that is restated as it is. */
int orc_cfmat_dist(const orc_asm *a, const orc_inv *inv, const char *cinvtype, double vpvs,
                   double freq, double py, const float complex *wave, float complex *fmat)
{
    const double twopi = 6.2831853071795862;
    int lisp = (cinvtype[0] == 'P' || cinvtype[0] == 'p') && (cinvtype[1] == 'P' || cinvtype[1] == 'p');
    int liss = (cinvtype[0] == 'S' || cinvtype[0] == 's') && (cinvtype[1] == 'S' || cinvtype[1] == 's');
    if (!lisp && !liss) return 1;
    if (inv->nvinv != 1) return 1;
    double omega = twopi * freq, vpvs2 = vpvs * vpvs;
    elem_ws w;
    ws_init(&w, a->nord);
    int q = w.nint * w.nint, nee = w.nee, nen = w.nen;
    double *dab = (double *)malloc(sizeof(double) * 2 * q);
    double *dvp = (double *)malloc(sizeof(double) * q), *dvs = (double *)malloc(sizeof(double) * q);
    float complex *forcev = (float complex *)malloc(sizeof(float complex) * (inv->mbufrhs + 1));
    int ierr = 0;
    for (int inpgl = 1; inpgl <= inv->nnpgl && !ierr; inpgl += inv->nvinv) {
        int inpg = -1;
        for (int g = 0; g < a->nnpg; g++)
            if ((inv->maskg[g] - 1) * inv->nvinv + 1 == inv->mygrad[inpgl - 1]) { inpg = g + 1; break; }
        if (inpg < 0) { ierr = 1; break; }
        int i1 = inv->jcsc[inpgl - 1], i2 = inv->jcsc[inpgl] - 1, nnz = i2 - i1 + 1;
        for (int i = i1; i <= i2; i++) fmat[i - 1] = 0.0f;
        for (int c = 0; c < inv->ncon && !ierr; c++) {
            int e = inv->mconn[(inpg - 1) + (size_t)a->nnpg * c];
            if (e <= 0) break;
            e--;
            if (orc_djacab(nen, w.nint, inpg, a->ieng + 4 * e, w.xigll, a->xlocs, a->zlocs, a->dens,
                           a->ecoeff, a->nnpg, w.shl, w.det, dab, w.shg)) { ierr = 1; break; }
            for (int k = 0; k < q; k++) {
                if (lisp) { dvp[k] = dab[k]; dvs[k] = dab[k] / vpvs2; }
                else { dvp[k] = dab[k + q] * vpvs2; dvs[k] = dab[k + q]; }
            }
            orc_dkm_vpvs(nen, w.nint, omega, py, w.wgll, dvp, dvs, w.det, w.shg, w.me, w.ke);
            for (int r = 0; r < nnz; r++) forcev[r] = 0.0f;       /* vforce25.f:252-256 */
            float complex cw = (float)inv->elem_wts[e] + I * 0.0f;
            const int *lm = a->lm + NDIM * nen * e;
            for (int ip = 0; ip < nee && !ierr; ip++) {
                int idof = lm[ip];
                if (idof <= 0) continue;
                int inz = ibsect(nnz, idof, inv->icsc + i1 - 1);
                if (inz <= 0) { ierr = 1; break; }
                for (int iq = 0; iq < nee; iq++) {
                    int jdof = lm[iq];
                    if (jdof <= 0) continue;
                    double complex v = w.me[ip + nee * iq] + w.ke[ip + nee * iq];
                    float complex carg = (float)creal(v) + I * (float)cimag(v);
                    forcev[inz - 1] += cmulf_ref(cmulf_ref(carg, wave[jdof - 1]), cw);
                }
            }
        }
        for (int r = 0; r < nnz; r++) fmat[i1 - 1 + r] = forcev[r];
    }
    free(dab); free(dvp); free(dvs); free(forcev);
    ws_free(&w);
    return ierr;
}

/* adjgrad.f90:88-145 ADJFLAM_DIST */
void orc_adjflam_dist(int nnpgl, int na35, const int *mygrad, const int *jcsc, const int *icsc,
                      const float complex *fmat, const float complex *adjlam, float *gprod)
{
    for (int i = 0; i < na35; i++) gprod[i] = 0.0f;
    for (int c = 0; c < nnpgl; c++) {
        int inpinv = mygrad[c];
        for (int i = jcsc[c]; i <= jcsc[c + 1] - 1; i++) {
            int jd = icsc[i - 1];
            float fr = crealf(fmat[i - 1]), fi = cimagf(fmat[i - 1]);
            float lr = crealf(adjlam[jd - 1]), li = cimagf(adjlam[jd - 1]);
            gprod[inpinv - 1] -= fr * lr + fi * li;
        }
    }
}

/* Anchor and inversion-variable index of local column inpgl (1-based), as
 * qpmfree.f90:781-789 and :1101-1109 find them (NVINV = 1). */
static int qp_anchor(const orc_asm *a, const orc_inv *inv, int inpgl, int *inpinv)
{
    for (int g = 0; g < a->nnpg; g++)
        if ((inv->maskg[g] - 1) * inv->nvinv + 1 == inv->mygrad[inpgl - 1]) {
            *inpinv = inv->maskg[g];
            return g + 1;
        }
    return -1;
}

/* qpmfree.f90:674-693, 1037-1055: PP -> P, SS -> S, PS -> both (NVINV = 2) */
static int qp_kind(const char *cinvtype, int *lisp, int *liss)
{
    char a0 = cinvtype[0] & ~0x20, a1 = cinvtype[1] & ~0x20;
    *lisp = (a0 == 'P' && a1 == 'P') || (a0 == 'P' && a1 == 'S');
    *liss = (a0 == 'S' && a1 == 'S') || (a0 == 'P' && a1 == 'S');
    return (*lisp || *liss) ? 0 : 1;
}

/* CARG(IVINV) for the element matrices of one element (qpmfree.f90:836-851,
 * 1163-1178): the P matrices first when LISP, then the S matrices */
static void qp_carg(int lisp, int liss, int nvinv, int at, const double *dmda, const double *dmdb,
                    const double complex *dkda, const double complex *dkdb, float complex *carg)
{
    int iv = 0;
    if (lisp) {
        double complex s = dmda[at] + dkda[at];
        carg[iv++] = (float)creal(s) + I * (float)cimag(s);
    }
    if (liss && iv < nvinv) {
        double complex s = dmdb[at] + dkdb[at];
        carg[iv++] = (float)creal(s) + I * (float)cimag(s);
    }
}

/* qpmfree.f90:622-912 CADJ_NU_ALPHA, its element loops: for every local
 * column (anchor INPG, variable IA35) and every element of MCONN(INPG,:),
 * with CWGHT = SNGL(ELEM_WTS(e)) * V(IA35) and CARG = CMPLX(DM + DK) from
 * DJACAB2 + DABM2KISOB25 (first derivatives; the second-derivative matrices
 * the reference also forms are not used by any output):
 *   RHSBNU(JDOF) += ADJLAM(IDOF) * CONJG(CWGHT*CARG)
 *   RHSALP(IDOF) += CWGHT*CARG*WAVE(JDOF)
 *   TEMP(row)    += CWGHT*CARG*WAVE(JDOF)   (cleared per element, :810-813)
 * and after the element loop P3(IA35) -= REAL(CONJG(TEMP(row)) ADJLAM(row))
 * over the column's rows.  The MUMPS solves and MPI reductions that follow
 * (:914-980) are the caller's.  COMPLEX*8 arithmetic in the reference's
 * order; PP or SS (NVINV = 1) or PS (NVINV = 2: per element entry the P
 * term, then the S term, each with the weight of its own variable). */
int orc_qp_nu_alpha(const orc_asm *a, const orc_inv *inv, const char *cinvtype, double freq,
                    double py, const float *v, const float complex *adjlam,
                    const float complex *wave, float complex *rhsbnu, float complex *rhsalp,
                    float *p3)
{
    const double twopi = 6.2831853071795862;
    int lisp, liss;
    if (qp_kind(cinvtype, &lisp, &liss) || inv->nvinv != lisp + liss) return 1;
    const int nv = inv->nvinv;
    double omega = twopi * freq;
    elem_ws w;
    ws_init(&w, a->nord);
    int q = w.nint * w.nint, nee = w.nee, nen = w.nen;
    double *dab = (double *)malloc(sizeof(double) * 2 * q), *dab2 = (double *)malloc(sizeof(double) * 2 * q);
    double *dmdb = (double *)malloc(sizeof(double) * nee * nee);
    double complex *dkdb = (double complex *)malloc(sizeof(double complex) * nee * nee);
    float complex *temp = (float complex *)malloc(sizeof(float complex) * 2 * (inv->mbufrhs + 1));
    for (int i = 0; i < a->ndof; i++) { rhsbnu[i] = 0.0f; rhsalp[i] = 0.0f; }
    for (int i = 0; i < inv->na35; i++) p3[i] = 0.0f;
    int ierr = 0;
    for (int inpgl = 1; inpgl <= inv->nnpgl && !ierr; inpgl += nv) {
        int inpinv, inpg = qp_anchor(a, inv, inpgl, &inpinv);
        if (inpg < 0) { ierr = 1; break; }
        int i1 = inv->jcsc[inpgl - 1], i2 = inv->jcsc[inpgl] - 1, nnz = i2 - i1 + 1;
        const int ld = inv->mbufrhs + 1;
        for (int r = 0; r < nv * ld; r++) temp[r] = 0.0f;
        for (int ic = 0; ic < inv->ncon && !ierr; ic++) {
            int e = inv->mconn[(inpg - 1) + (size_t)a->nnpg * ic];
            if (e <= 0) break;
            e--;
            if (orc_djacab2(nen, w.nint, inpg, a->ieng + 4 * e, w.xigll, a->xlocs, a->zlocs, a->dens,
                            a->ecoeff, a->nnpg, w.shl, w.det, dab, dab2, w.shg)) { ierr = 1; break; }
            orc_dabmkisob25(nen, w.nint, lisp, liss, omega, py, w.wgll, w.det, dab, w.shg,
                            w.me, dmdb, w.ke, dkdb);
            for (int r = 0; r < nv * ld; r++) temp[r] = 0.0f;
            const int *lm = a->lm + NDIM * nen * e;
            for (int ip = 0; ip < nee && !ierr; ip++) {
                int idof = lm[ip];
                if (idof <= 0) continue;
                int inz = ibsect(nnz, idof, inv->icsc + i1 - 1);
                if (inz <= 0) { ierr = 1; break; }
                for (int iq = 0; iq < nee; iq++) {
                    int jdof = lm[iq];
                    if (jdof <= 0) continue;
                    float complex carg[2];
                    qp_carg(lisp, liss, nv, ip + nee * iq, w.me, dmdb, w.ke, dkdb, carg);
                    for (int iv = 0; iv < nv; iv++) {
                        int ia35 = (inpinv - 1) * nv + iv + 1;
                        float complex cw = (float)inv->elem_wts[e] * v[ia35 - 1] + I * 0.0f;
                        float complex wc = cmulf_ref(cw, carg[iv]);
                        rhsbnu[jdof - 1] += cmulf_ref(adjlam[idof - 1], conjf(wc));
                        rhsalp[idof - 1] += cmulf_ref(wc, wave[jdof - 1]);
                        temp[iv * ld + inz - 1] += cmulf_ref(wc, wave[jdof - 1]);
                    }
                }
            }
        }
        for (int iv = 0; iv < nv && !ierr; iv++) {
            int ia35 = (inpinv - 1) * nv + iv + 1;
            int k1 = inv->jcsc[inpgl + iv - 1], k2 = inv->jcsc[inpgl + iv] - 1;
            for (int r = 0; r <= k2 - k1; r++) {
                int idof = inv->icsc[k1 - 1 + r];
                p3[ia35 - 1] -= crealf(cmulf_ref(conjf(temp[iv * ld + r]), adjlam[idof - 1]));
            }
        }
    }
    free(dab); free(dab2); free(dmdb); free(dkdb); free(temp);
    ws_free(&w);
    return ierr;
}

/* qpmfree.f90:1000-1215 MET225_MOD, its element loops: per local column the
 * elements of MCONN(INPG,:) with CWGHT = SNGL(ELEM_WTS(e)) and CARG from
 * DJACAB + DABMKISOB25,
 *   T1(row) += CWGHT*CARG*WAVE(JDOF),  T2(row) += CWGHT*CARG*ALPHA(JDOF)
 * (both cleared per element, :1155-1160, so the last element remains), then
 * P1(IA35) -= REAL(CONJG(T1) NU), P2(IA35) -= REAL(CONJG(T2) ADJLAM) over
 * the column's rows; before the MPI reductions of :1222-1226.  PS (NVINV =
 * 2): T1, T2 per variable from the P and the S matrices of the element. */
int orc_met225(const orc_asm *a, const orc_inv *inv, const char *cinvtype, double freq, double py,
               const float complex *adjlam, const float complex *nu, const float complex *alpha,
               const float complex *wave, float *p1, float *p2)
{
    const double twopi = 6.2831853071795862;
    int lisp, liss;
    if (qp_kind(cinvtype, &lisp, &liss) || inv->nvinv != lisp + liss) return 1;
    const int nv = inv->nvinv;
    double omega = twopi * freq;
    elem_ws w;
    ws_init(&w, a->nord);
    int q = w.nint * w.nint, nee = w.nee, nen = w.nen;
    double *dab = (double *)malloc(sizeof(double) * 2 * q);
    double *dmdb = (double *)malloc(sizeof(double) * nee * nee);
    double complex *dkdb = (double complex *)malloc(sizeof(double complex) * nee * nee);
    const int ld = inv->mbufrhs + 1;
    float complex *t1 = (float complex *)malloc(sizeof(float complex) * 2 * ld);
    float complex *t2 = (float complex *)malloc(sizeof(float complex) * 2 * ld);
    for (int i = 0; i < inv->na35; i++) { p1[i] = 0.0f; p2[i] = 0.0f; }
    int ierr = 0;
    for (int inpgl = 1; inpgl <= inv->nnpgl && !ierr; inpgl += nv) {
        int inpinv, inpg = qp_anchor(a, inv, inpgl, &inpinv);
        if (inpg < 0) { ierr = 1; break; }
        int i1 = inv->jcsc[inpgl - 1], i2 = inv->jcsc[inpgl] - 1, nnz = i2 - i1 + 1;
        for (int r = 0; r < nv * ld; r++) { t1[r] = 0.0f; t2[r] = 0.0f; }
        for (int ic = 0; ic < inv->ncon && !ierr; ic++) {
            int e = inv->mconn[(inpg - 1) + (size_t)a->nnpg * ic];
            if (e <= 0) break;
            e--;
            float complex cw = (float)inv->elem_wts[e] + I * 0.0f;
            if (orc_djacab(nen, w.nint, inpg, a->ieng + 4 * e, w.xigll, a->xlocs, a->zlocs, a->dens,
                           a->ecoeff, a->nnpg, w.shl, w.det, dab, w.shg)) { ierr = 1; break; }
            orc_dabmkisob25(nen, w.nint, lisp, liss, omega, py, w.wgll, w.det, dab, w.shg,
                            w.me, dmdb, w.ke, dkdb);
            for (int r = 0; r < nv * ld; r++) { t1[r] = 0.0f; t2[r] = 0.0f; }
            const int *lm = a->lm + NDIM * nen * e;
            for (int ip = 0; ip < nee && !ierr; ip++) {
                int idof = lm[ip];
                if (idof <= 0) continue;
                int inz = ibsect(nnz, idof, inv->icsc + i1 - 1);
                if (inz <= 0) { ierr = 1; break; }
                for (int iq = 0; iq < nee; iq++) {
                    int jdof = lm[iq];
                    if (jdof <= 0) continue;
                    float complex carg[2];
                    qp_carg(lisp, liss, nv, ip + nee * iq, w.me, dmdb, w.ke, dkdb, carg);
                    for (int iv = 0; iv < nv; iv++) {
                        float complex wc = cmulf_ref(cw, carg[iv]);
                        t1[iv * ld + inz - 1] += cmulf_ref(wc, wave[jdof - 1]);
                        t2[iv * ld + inz - 1] += cmulf_ref(wc, alpha[jdof - 1]);
                    }
                }
            }
        }
        for (int iv = 0; iv < nv && !ierr; iv++) {
            int ia35 = (inpinv - 1) * nv + iv + 1;
            int k1 = inv->jcsc[inpgl + iv - 1], k2 = inv->jcsc[inpgl + iv] - 1;
            for (int r = 0; r <= k2 - k1; r++) {
                int idof = inv->icsc[k1 - 1 + r];
                p1[ia35 - 1] -= crealf(cmulf_ref(conjf(t1[iv * ld + r]), nu[idof - 1]));
                p2[ia35 - 1] -= crealf(cmulf_ref(conjf(t2[iv * ld + r]), adjlam[idof - 1]));
            }
        }
    }
    free(dab); free(dmdb); free(dkdb); free(t1); free(t2);
    ws_free(&w);
    return ierr;
}

/* disp25.f90:710-908 FILL_UWAVE_QMESH (LCHECK false): the plane wave
 * u = (cos BAZ sin AOI, sin BAZ sin AOI, cos AOI) exp(-i (KAPPAX x + KAPPAZ z))
 * at every DOF, from the first element (in element, eta, xi order) holding
 * it; x, z by CSF2DND at the GLL point.  Returns 1 when a DOF is never
 * reached. */
int orc_fill_uwave(const orc_asm *a, double kappax, double kappaz, double baz, double aoi,
                   double complex *ue)
{
    const double pi180 = 0.017453292519943295;
    const int nl = a->nord + 1, nen = nl * nl;
    double *xi = (double *)malloc(sizeof(double) * nl), *wt = (double *)malloc(sizeof(double) * nl);
    char *init = (char *)calloc(a->ndof > 0 ? a->ndof : 1, 1);
    orc_gll(nl, xi, wt);
    if (nl % 2) xi[(nl - 1) / 2] = 0.0;   /* GENPTS, dshl.f:395-396 */
    const double cosaz = cos(baz * pi180), sinaz = sin(baz * pi180);
    const double cosi = cos(aoi * pi180), sini = sin(aoi * pi180);
    const double pu = cosaz * sini, pv = sinaz * sini, pw = cosi;
    for (int i = 0; i < a->ndof; i++) ue[i] = 0.0;
    for (int e = 0; e < a->nelem; e++)
        for (int ie = 0; ie < nl; ie++)
            for (int ix = 0; ix < nl; ix++) {
                const int iae = ie * nl + ix;
                const int *lm = a->lm + NDIM * (iae + nen * e);
                int seen = 0;
                for (int i = 0; i < NDIM; i++)
                    if (lm[i] > 0 && init[lm[i] - 1]) seen = 1;
                if (seen) continue;
                double sf[4], x = 0.0, z = 0.0;
                csf2dnd(xi[ix], xi[ie], sf);
                for (int k = 0; k < 4; k++) {
                    int g = a->ieng[4 * e + k] - 1;
                    x += sf[k] * a->xlocs[g];
                    z += sf[k] * a->zlocs[g];
                }
                double complex carg = cexp(CMPLX(0.0, -kappax * x - kappaz * z));
                const double p[3] = {pu, pv, pw};
                for (int i = 0; i < NDIM; i++) {
                    int d = lm[i];
                    if (d <= 0) continue;
                    init[d - 1] = 1;
                    ue[d - 1] = CMPLX(p[i] * creal(carg), p[i] * cimag(carg));
                }
            }
    int ierr = 0;
    for (int i = 0; i < a->ndof; i++)
        if (!init[i]) ierr = 1;
    free(xi); free(wt); free(init);
    return ierr;
}

/* ------------------------------------------------------------------------ */
/* Adjoint-source pipeline: receiver extraction, residuals, adjoint RHS.
 * COMPLEX*8 / REAL*4 arithmetic as the reference compiles it: products by
 * the textbook formula, quotients through the runtime's complex divide,
 * elementary functions from libm's single-precision routines. */

/* locrec2d.f:867-925 EXRESP25D: EST(i,irec,isrc) = RECV(i,irec) *
 * SOL((isrc-1)*NDOF + MRDOF(i,irec)) * exp(+i w PY(isrc) YREC(irec)), zero
 * where MRDOF is 0; the phase argument is rounded to REAL*4 first (CMPLX of
 * default kind).  MDIM = 3, MREC = NREC. */
void orc_exresp25d(int ndof, int nsrc, int nrec, double freq, const double *py, const double *yrec,
                   const float complex *sol, const int *mrdof, const float complex *recv,
                   float complex *est)
{
    const double twopi = 6.2831853071795862;
    double omega = twopi * freq;
    for (int is = 0; is < nsrc; is++) {
        long loff = (long)is * ndof;
        for (int ir = 0; ir < nrec; ir++) {
            float complex cfact = cexpf(CMPLXF(0.0f, (float)(omega * py[is] * yrec[ir])));
            for (int i = 0; i < NDIM; i++) {
                float complex *o = est + NDIM * ((long)nrec * is + ir) + i;
                int d = mrdof[NDIM * ir + i];
                if (d != 0) *o = cmulf_ref(cmulf_ref(recv[NDIM * ir + i], sol[loff + d - 1]), cfact);
                else *o = 0.0f;
            }
        }
    }
}

/* wrapph.f:178-191 CROTATE (angle in degrees, REAL*4) */
void orc_crotate(float theta, float complex q1, float complex q2, float complex *rq1, float complex *rq2)
{
    const float pi180 = 0.017453292519943295f;
    float complex cost = CMPLXF(cosf(theta * pi180), 0.0f), sint = CMPLXF(sinf(theta * pi180), 0.0f);
    float complex a = cmulf_ref(q1, cost), b = cmulf_ref(q2, sint);
    float complex c2 = cmulf_ref(q1, sint), d = cmulf_ref(q2, cost);
    *rq1 = CMPLXF(crealf(a) - crealf(b), cimagf(a) - cimagf(b));
    *rq2 = CMPLXF(crealf(c2) + crealf(d), cimagf(c2) + cimagf(d));
}

/* wrapph.f:169-174 CPHM2CM: magnitude, phase -> complex */
float complex orc_cphm2cm(float xmag, float phase)
{
    return cmulf_ref(CMPLXF(xmag, 0.0f), cexpf(CMPLXF(0.0f, phase)));
}

static float sphase(float complex q) { return atan2f(cimagf(q), crealf(q)); }   /* wrapph.f:154-159 */

/* COMPLEX*8 quotient as the reference's compiler forms it: the textbook
 * formula in single precision, without scaling (matches the compiled
 * BPRESID4 bit for bit; a scaled or promoted divide does not) */
static inline float complex cdivf_ref(float complex a, float complex b)
{
    float ar = crealf(a), ai = cimagf(a), br = crealf(b), bi = cimagf(b);
    float d = br * br + bi * bi;
    return CMPLXF((ar * br + ai * bi) / d, (ai * br - ar * bi) / d);
}

/* objfn.f:972-1009 CRESID: 1 phase only, 2 amplitude only, else OBS - EST */
float complex orc_cresid(int irestp, float complex obs, float complex est)
{
    if (crealf(obs) == 0.0f && cimagf(obs) == 0.0f) return 0.0f;
    if (irestp == 1 || irestp == 2) {
        float robs = crealf(obs), zobs = cimagf(obs), rest = crealf(est), zest = cimagf(est);
        float adata, pdata, aest, pest;
        if (irestp == 2) {
            adata = sqrtf(robs * robs + zobs * zobs); pdata = 0.0f;
            aest = sqrtf(rest * rest + zest * zest); pest = 0.0f;
        } else {
            adata = 1.0f; pdata = atan2f(zobs, robs);
            aest = 1.0f; pest = atan2f(zest, rest);
        }
        float complex a = cmulf_ref(CMPLXF(adata, 0.0f), cexpf(CMPLXF(0.0f, pdata)));
        float complex b = cmulf_ref(CMPLXF(aest, 0.0f), cexpf(CMPLXF(0.0f, pest)));
        return CMPLXF(crealf(a) - crealf(b), cimagf(a) - cimagf(b));
    }
    return CMPLXF(crealf(obs) - crealf(est), cimagf(obs) - cimagf(est));
}

/* objfn.f:757-839 BPRESID4 (its EPS is 0, so the Huber and combination
 * branches give what those formulas give with EPS = 0) */
void orc_bpresid4(int nrec, int norm, int irestp, float epss, double freq, float dtmax,
                  const float *swght, const float complex *obs, const float complex *est,
                  float complex *resid)
{
    const float twopi = 6.283185307179586f, pi = 3.1415926535897931f;
    float eps = 0.0f, eps2 = epss * (eps * eps);
    for (int ir = 0; ir < nrec; ir++)
        for (int i = 0; i < NDIM; i++) {
            int k = NDIM * ir + i;
            resid[k] = 0.0f;
            if (cabsf(obs[k]) == 0.0f) continue;
            float complex dif = orc_cresid(irestp, obs[k], est[k]);
            float complex weight = CMPLXF(swght[k], 0.0f);
            if (norm == 1) {
                resid[k] = cdivf_ref(cmulf_ref(conjf(weight), dif), CMPLXF(cabsf(dif), 0.0f));
            } else if (norm == 3) {
                if (cabsf(cmulf_ref(weight, dif)) <= eps) {
                    float w2 = cabsf(weight) * cabsf(weight) / eps;
                    resid[k] = cmulf_ref(CMPLXF(w2, 0.0f), dif);
                } else {
                    resid[k] = cdivf_ref(cmulf_ref(conjf(weight), dif), CMPLXF(cabsf(dif), 0.0f));
                }
            } else if (norm == 4) {
                float w2 = cabsf(weight) * cabsf(weight);
                float a = cabsf(cmulf_ref(weight, dif));
                float denon = eps2 * sqrtf(1.0f + a * a / eps2);
                resid[k] = cmulf_ref(CMPLXF(w2 / denon, 0.0f), dif);
            } else {
                float w2 = cabsf(weight) * cabsf(weight);
                resid[k] = cmulf_ref(CMPLXF(w2, 0.0f), dif);
                if (dtmax > 0.0f) {
                    float dphi = sphase(obs[k]) - sphase(est[k]);
                    if (dphi < -pi) dphi = dphi + twopi;
                    if (dphi > +pi) dphi = dphi - twopi;
                    if (fabs((double)dphi / ((double)twopi * freq)) > (double)dtmax) resid[k] = 0.0f;
                }
            }
        }
}

/* adjgrad.f90:527-566 HDIFFER: spreading compensation factor */
float complex orc_hdiffer(int ibphase, double freq)
{
    const double pi = 3.1415926535897932384626434;
    float omega = (float)(2.0 * pi * freq);
    if (ibphase == 1 || ibphase == 2) {
        float complex h = csqrtf(CMPLXF(0.0f, -omega));
        return ibphase == 2 ? -h : h;
    }
    if (ibphase == 3 || ibphase == 4) {
        float complex h = csqrtf(CMPLXF(0.0f, omega));
        return ibphase == 4 ? -h : h;
    }
    if (ibphase == 5) return CMPLXF(0.0f, omega);
    if (ibphase == 6) return CMPLXF(0.0f, 1.0f);
    return CMPLXF(1.0f, 0.0f);
}

/* adjgrad.f90:448-523 RESID_DRHS: RHS(MRDOF(i,irec)) += HFACT*RESID(i,irec)
 * in receiver order, invalid DOFs skipped with IWARN = 1, then RHS = CONJG */
void orc_resid_drhs(int ndof, int nrec, int ibphase, double freq, const int *mrdof,
                    const float complex *resid, float complex *rhs, int *iwarn)
{
    float complex h = orc_hdiffer(ibphase, freq);
    *iwarn = 0;
    for (int i = 0; i < ndof; i++) rhs[i] = 0.0f;
    for (int ir = 0; ir < nrec; ir++)
        for (int i = 0; i < NDIM; i++) {
            int d = mrdof[NDIM * ir + i];
            if (d <= 0 || d > ndof) { *iwarn = 1; continue; }
            float complex t = cmulf_ref(h, resid[NDIM * ir + i]);
            rhs[d - 1] = CMPLXF(crealf(rhs[d - 1]) + crealf(t), cimagf(rhs[d - 1]) + cimagf(t));
        }
    for (int i = 0; i < ndof; i++) rhs[i] = conjf(rhs[i]);
}

/* element-wise forms of the scalar routines above, for callers that cannot
 * receive a complex return value */
void orc_cresid_n(int n, int irestp, const float complex *obs, const float complex *est, float complex *out)
{
    for (int i = 0; i < n; i++) out[i] = orc_cresid(irestp, obs[i], est[i]);
}

void orc_cphm2cm_n(int n, const float *xmag, const float *ph, float complex *out)
{
    for (int i = 0; i < n; i++) out[i] = orc_cphm2cm(xmag[i], ph[i]);
}

void orc_crotate_n(int n, float theta, const float complex *q1, const float complex *q2, float complex *r1,
                   float complex *r2)
{
    for (int i = 0; i < n; i++) orc_crotate(theta, q1[i], q2[i], &r1[i], &r2[i]);
}

void orc_hdiffer_v(int ibphase, double freq, float *out)
{
    float complex h = orc_hdiffer(ibphase, freq);
    out[0] = crealf(h);
    out[1] = cimagf(h);
}

/* adjgrad.f90:240-293, the loop that readies receiver irec of (IFREQ, ISRC)
 * for the residual: estimates rotated to (N,E,Z) by AZMOD (magnitude/phase
 * pairs first turned complex when LUNWRAP), data and estimates shifted by
 * exp(-i SNGL(2 pi FREQ PY YREC)).  Fills o, e, w (3 x nrec) and, when raw
 * is given, the untransformed observations inv%OBS(1:3, IFREQ, :, ISRC). */
static void cadj_frames(int nfreq, int nrec, int ifreq, int isrc, double freq, double py, double azmod,
                        int lunwrap, const double *yrec, const float complex *est, const float complex *obs,
                        const float *wghts, float complex *o, float complex *e, float *w, float complex *raw)
{
    const double twopi = 6.2831853071795862;
    for (int ir = 0; ir < nrec; ir++) {
        long base = NDIM * ((long)(ifreq - 1) + (long)nfreq * (ir + (long)nrec * (isrc - 1)));
        const float complex *ev = est + base, *ov = obs + base;
        for (int i = 0; i < NDIM; i++) w[NDIM * ir + i] = wghts[base + i];
        if (raw)
            for (int i = 0; i < NDIM; i++) raw[NDIM * ir + i] = ov[i];
        float complex q1, q2, e3, o1, o2, o3;
        if (lunwrap) {
            q1 = orc_cphm2cm(crealf(ev[0]), cimagf(ev[0]));
            q2 = orc_cphm2cm(crealf(ev[1]), cimagf(ev[1]));
            e3 = orc_cphm2cm(crealf(ev[2]), cimagf(ev[2]));
            o1 = orc_cphm2cm(crealf(ov[0]), cimagf(ov[0]));
            o2 = orc_cphm2cm(crealf(ov[1]), cimagf(ov[1]));
            o3 = orc_cphm2cm(crealf(ov[2]), cimagf(ov[2]));
        } else {
            q1 = ev[0]; q2 = ev[1]; e3 = ev[2];
            o1 = ov[0]; o2 = ov[1]; o3 = ov[2];
        }
        float complex *E = e + NDIM * ir, *O = o + NDIM * ir;
        orc_crotate((float)azmod, q1, q2, &E[0], &E[1]);
        E[2] = e3;
        O[0] = o1; O[1] = o2; O[2] = o3;
        float complex cfact = cexpf(CMPLXF(0.0f, -(float)(twopi * freq * py * yrec[ir])));
        for (int i = 0; i < NDIM; i++) {
            O[i] = cmulf_ref(O[i], cfact);
            E[i] = cmulf_ref(E[i], cfact);
        }
    }
}

/* adjgrad.f90:149-300 CADJLAM up to the solve (no data covariance, LCOVD
 * false): cadj_frames, BPRESID4, residuals rotated back by -AZMOD,
 * RESID_DRHS.  est, obs: (3, nfreq, nrec, nsrc) COMPLEX*8; wghts REAL*4
 * likewise; ifreq, isrc 1-based. */
void orc_cadjlam_rhs(int ndof, int nfreq, int nrec, int ifreq, int isrc, double freq, double py,
                     double azmod, int norm, int irestp, float epss, float dtmax, int ibphase,
                     int lunwrap, const double *yrec, const int *mrdof, const float complex *est,
                     const float complex *obs, const float *wghts, float complex *rhs, int *iwarn)
{
    size_t n3 = (size_t)NDIM * (nrec > 0 ? nrec : 1);
    float complex *e = (float complex *)calloc(n3, sizeof(float complex));
    float complex *o = (float complex *)calloc(n3, sizeof(float complex));
    float complex *r = (float complex *)calloc(n3, sizeof(float complex));
    float *w = (float *)calloc(n3, sizeof(float));
    cadj_frames(nfreq, nrec, ifreq, isrc, freq, py, azmod, lunwrap, yrec, est, obs, wghts, o, e, w, NULL);
    orc_bpresid4(nrec, norm, irestp, epss, freq, dtmax, w, o, e, r);
    for (int ir = 0; ir < nrec; ir++) {
        float complex *R = r + NDIM * ir, qn = R[0], qe = R[1];
        orc_crotate(-(float)azmod, qn, qe, &R[0], &R[1]);
    }
    orc_resid_drhs(ndof, nrec, ibphase, freq, mrdof, r, rhs, iwarn);
    free(e); free(o); free(r); free(w);
}

/* locrec2d.f:157-223 REC_SPACE1D: dx(i) = |x(i+1) - x(i)| / min_j |x(j+1) -
 * x(j)|.  Returns 1 with dx zeroed for a single receiver, a spacing below
 * 1e-10, or a reversal of direction (checked, as the reference checks it,
 * at receivers 2..nrec-1 against the first spacing's sign). */
int orc_rec_space1d(int nrec, const double *xrec, double *dx)
{
    for (int i = 0; i < nrec - 1; i++) dx[i] = 0.0;
    if (nrec == 1) return 1;
    double xdir = xrec[1] - xrec[0], dxmin = HUGE_VAL;
    for (int i = 0; i < nrec - 1; i++) {          /* i = IREC - 1 */
        dx[i] = fabs(xrec[i + 1] - xrec[i]);
        dxmin = fmin(dxmin, dx[i]);
        if (dx[i] < 1e-10 ||
            (i > 0 && ((xrec[i] > xrec[i - 1] && xdir < 0.0) || (xrec[i] < xrec[i - 1] && xdir > 0.0)))) {
            for (int j = 0; j < nrec - 1; j++) dx[j] = 0.0;
            return 1;
        }
    }
    for (int i = 0; i < nrec - 1; i++) dx[i] = dx[i] / dxmin;
    return 0;
}

/* locrec2d.f:334-509 LAPLACE_1DV2: the (6 nrec x 6 nrec, column-major) data
 * roughening matrix.  Components are taken in the order 3, 2, 1; the
 * receivers with a non-zero observation of a component are the nodes of a
 * chain of two-node elements with stiffness (2/h)(1/2)[1 -1; -1 1], h the
 * summed normalised spacing to the next such receiver (1 without ldwght);
 * the nodes take consecutive numbers over all components (a component with
 * no observation takes none), and the same block is added at row/column
 * offset 3 nrec for the imaginary parts.  A zero observation matrix (dead
 * frequency) leaves COVD zero; returns 1 when a spacing sums to zero. */
int orc_laplace_1dv2(int nrec, int ldwght, const double *dx, const float complex *obs, double *covd)
{
    const int ldc = 2 * NDIM * nrec;
    memset(covd, 0, sizeof(double) * (size_t)ldc * ldc);
    float amax = 0.0f;
    for (int k = 0; k < NDIM * nrec; k++) amax = fmaxf(amax, cabsf(obs[k]));
    if (amax == 0.0f) return 0;
    int *act = (int *)malloc(sizeof(int) * nrec);
    double *h = (double *)malloc(sizeof(double) * nrec);
    int idof = 0, ierr = 0;
    for (int c = NDIM - 1; c >= 0 && !ierr; c--) {
        int na = 0;
        for (int ir = 0; ir < nrec; ir++)
            if (cabsf(obs[NDIM * ir + c]) > 0.0f) act[na++] = ir;
        if (na == 0) continue;                         /* dead component */
        const int nel = na - 1;
        for (int j = 0; j < nel && !ierr; j++) {
            if (ldwght) {
                double xs = 0.0;
                for (int r = act[j]; r < act[j + 1]; r++) xs += dx[r];
                h[j] = xs;
                if (xs == 0.0) ierr = 1;
            } else {
                h[j] = 1.0;
            }
        }
        if (ierr) break;
        for (int j = 0; j < nel; j++) {
            const int d[2] = {idof + j, idof + j + 1};   /* 0-based node numbers */
            const double kk[2][2] = {{0.5, -0.5}, {-0.5, 0.5}};
            for (int a = 0; a < 2; a++)
                for (int b = 0; b < 2; b++) {
                    const double t = 2.0 / h[j] * kk[a][b];
                    covd[(size_t)d[b] * ldc + d[a]] += t;
                    covd[(size_t)(NDIM * nrec + d[b]) * ldc + NDIM * nrec + d[a]] += t;
                }
        }
        idof += na;
    }
    free(act); free(h);
    return ierr;
}

/* locrec2d.f:511-549 PERM_OBS: iperm[3 irec + i] (1-based positions) lists,
 * for components 3, 2, 1 in turn, the receivers with a non-zero observation
 * and then those without. */
int orc_perm_obs(int nrec, const float complex *obs, int *iperm)
{
    int iobs = 0;
    for (int k = 0; k < NDIM * nrec; k++) iperm[k] = 0;
    for (int c = NDIM - 1; c >= 0; c--) {
        for (int ir = 0; ir < nrec; ir++)
            if (cabsf(obs[NDIM * ir + c]) > 0.0f) iperm[NDIM * ir + c] = ++iobs;
        for (int ir = 0; ir < nrec; ir++)
            if (cabsf(obs[NDIM * ir + c]) == 0.0f) iperm[NDIM * ir + c] = ++iobs;
    }
    for (int k = 0; k < NDIM * nrec; k++)
        if (iperm[k] == 0) return 1;
    return 0;
}

/* objfn.f:635-733 BPRESID_COVD4: WORK holds CRESID(IRESTP) of every
 * receiver component with a non-zero observation at its permuted position
 * (real parts, then imaginary parts 3 nrec further on), zeroed where the
 * phase gate (DTMAX > 0, in seconds, single precision) fails; COVD = SNGL
 * of COVD8 scaled on both sides by the receiver weights at the permuted
 * positions; RESIDV = COVD WORK as SGEMV forms it (column by column, single
 * precision); RESID(i, irec) from RESIDV at the receiver's positions. */
void orc_bpresid_covd4(int nrec, const int *iperm, int irestp, float dtmax, double freq,
                       const double *covd8, const float *swght, const float complex *obs,
                       const float complex *est, float complex *resid)
{
    const float twopi = 6.283185307179586f, pi = 3.1415926535897931f;
    const int n3 = NDIM * nrec, m = 2 * n3;
    float *work = (float *)calloc((size_t)m, sizeof(float));
    float *wp = (float *)calloc((size_t)m, sizeof(float));
    float *rv = (float *)calloc((size_t)m, sizeof(float));
    for (int k = 0; k < n3; k++) {
        const int l = iperm[k] - 1;
        wp[l] = swght[k];
        wp[n3 + l] = swght[k];
        if (cabsf(obs[k]) > 0.0f) {
            const float complex dif = orc_cresid(irestp, obs[k], est[k]);
            work[l] = crealf(dif);
            work[n3 + l] = cimagf(dif);
            if (dtmax > 0.0f) {
                float dphi = sphase(obs[k]) - sphase(est[k]);
                if (dphi < -pi) dphi = dphi + twopi;
                if (dphi > +pi) dphi = dphi - twopi;
                dphi = dphi / (twopi * (float)freq);
                if (fabsf(dphi) > dtmax) {
                    work[l] = 0.0f;
                    work[n3 + l] = 0.0f;
                }
            }
        }
    }
    for (int j = 0; j < m; j++) {                       /* SGEMV 'N', beta = 0 */
        const float t = work[j];
        for (int i = 0; i < m; i++) {
            const float a = (wp[i] * (float)covd8[(size_t)j * m + i]) * wp[j];
            rv[i] = rv[i] + t * a;
        }
    }
    for (int k = 0; k < n3; k++) {
        const int l = iperm[k] - 1;
        resid[k] = CMPLXF(rv[l], rv[n3 + l]);
    }
    free(work); free(wp); free(rv);
}

/* adjgrad.f90:149-300 CADJLAM up to the solve with the data covariance
 * (LCOVD): LAPLACE_1DV2 and PERM_OBS of the untransformed observations,
 * cadj_frames, BPRESID_COVD4, residuals rotated back by -AZMOD,
 * RESID_DRHS.  Arguments as orc_cadjlam_rhs, with ldwght and dx (nrec - 1
 * normalised spacings, REC_SPACE1D) in place of norm and epss.  Returns
 * LAPLACE_1DV2's or PERM_OBS's error flag (rhs untouched then). */
int orc_cadjlam_rhs_covd(int ndof, int nfreq, int nrec, int ifreq, int isrc, double freq, double py,
                         double azmod, int irestp, float dtmax, int ibphase, int lunwrap, int ldwght,
                         const double *dx, const double *yrec, const int *mrdof, const float complex *est,
                         const float complex *obs, const float *wghts, float complex *rhs, int *iwarn)
{
    const size_t n3 = (size_t)NDIM * (nrec > 0 ? nrec : 1), ldc = 2 * n3;
    float complex *e = (float complex *)calloc(n3, sizeof(float complex));
    float complex *o = (float complex *)calloc(n3, sizeof(float complex));
    float complex *raw = (float complex *)calloc(n3, sizeof(float complex));
    float complex *r = (float complex *)calloc(n3, sizeof(float complex));
    float *w = (float *)calloc(n3, sizeof(float));
    int *iperm = (int *)calloc(n3, sizeof(int));
    double *covd = (double *)malloc(sizeof(double) * ldc * ldc);
    cadj_frames(nfreq, nrec, ifreq, isrc, freq, py, azmod, lunwrap, yrec, est, obs, wghts, o, e, w, raw);
    int ierr = orc_laplace_1dv2(nrec, ldwght, dx, raw, covd);
    if (!ierr) ierr = orc_perm_obs(nrec, raw, iperm);
    if (!ierr) {
        orc_bpresid_covd4(nrec, iperm, irestp, dtmax, freq, covd, w, o, e, r);
        for (int ir = 0; ir < nrec; ir++) {
            float complex *R = r + NDIM * ir, qn = R[0], qe = R[1];
            orc_crotate(-(float)azmod, qn, qe, &R[0], &R[1]);
        }
        orc_resid_drhs(ndof, nrec, ibphase, freq, mrdof, r, rhs, iwarn);
    }
    free(e); free(o); free(raw); free(r); free(w); free(iperm); free(covd);
    return ierr;
}

/* Exact derivative of S(w,ky) with respect to anchor lambda/mu/rho.  The
 * element matrix of KM25ISOQ is linear in the interpolated (lambda, mu, rho)
 * fields, so dS_e/d(lambda_k) is KM25ISOQ evaluated with lambda(q) = N_k(q),
 * mu = rho = 0 (and likewise for mu, rho).  PML elements are excluded. */
int orc_grad_exact(const orc_asm *a, double freq, double py, const double complex *u,
                   const double complex *lam, double *g_lmr, double *g_vvr)
{
    const double twopi = 6.2831853071795862;
    double omega = twopi * freq;
    elem_ws w;
    ws_init(&w, a->nord);
    int q = w.nint * w.nint, nee = w.nee, nen = w.nen;
    double *fl = (double *)malloc(sizeof(double) * q), *fm = (double *)malloc(sizeof(double) * q);
    double *fr = (double *)malloc(sizeof(double) * q);
    double *nk = (double *)malloc(sizeof(double) * 4 * q);
    int ierr = 0;
    for (int e = 0; e < a->nelem && !ierr; e++) {
        if (a->cdomain[e] == 'A') continue;
        const int *ie = a->ieng + 4 * e;
        if (orc_cjmk_tvi(nen, w.nint, ie, w.xigll, a->xlocs, a->zlocs, a->dens, a->ecoeff,
                         a->nnpg, w.shl, w.det, w.rho, w.dmat, w.shg)) { ierr = 1; break; }
        for (int ix = 0; ix < w.nint; ix++)
            for (int iz = 0; iz < w.nint; iz++) {
                double sf[4], dsf[8];
                orc_csf2d(w.xigll[ix], w.xigll[iz], sf, dsf);
                for (int k = 0; k < 4; k++) nk[k + 4 * (ix + w.nint * iz)] = sf[k];
            }
        const int *lm = a->lm + NDIM * nen * e;
        for (int k = 0; k < 4; k++)
            for (int p = 0; p < 3; p++) {
                for (int j = 0; j < q; j++) {
                    fl[j] = (p == 0) ? nk[k + 4 * j] : 0.0;
                    fm[j] = (p == 1) ? nk[k + 4 * j] : 0.0;
                    fr[j] = (p == 2) ? nk[k + 4 * j] : 0.0;
                }
                orc_km25isoq(nen, w.nint, omega, py, w.wgll, fl, fm, fr, w.det, w.shg, w.me, w.ke);
                double complex acc = 0.0;
                for (int ip = 0; ip < nee; ip++) {
                    int id = lm[ip];
                    if (id <= 0) continue;
                    double complex row = 0.0;
                    for (int iq = 0; iq < nee; iq++) {
                        int jd = lm[iq];
                        if (jd <= 0) continue;
                        row += (w.me[ip + nee * iq] + w.ke[ip + nee * iq]) * u[jd - 1];
                    }
                    acc += conj(lam[id - 1]) * row;
                }
                g_lmr[3 * (ie[k] - 1) + p] += creal(acc);
            }
    }
    /* anchors of PML elements hold fixed parameters (the reference inverts
     * no anchor of an absorbing element, gengmask.f90:110-121): no entries */
    for (int e = 0; e < a->nelem; e++)
        if (a->cdomain[e] == 'A')
            for (int k = 0; k < 4; k++)
                for (int p = 0; p < 3; p++) g_lmr[3 * (a->ieng[4 * e + k] - 1) + p] = 0.0;
    if (g_vvr) {
        for (int g = 0; g < a->nnpg; g++) {
            double l = a->ecoeff[g], m = a->ecoeff[g + a->nnpg], r = a->dens[g];
            double vp = sqrt((l + 2.0 * m) / r), vs = sqrt(m / r);
            double gl = g_lmr[3 * g], gm = g_lmr[3 * g + 1], gr = g_lmr[3 * g + 2];
            g_vvr[3 * g + 0] = gl * 2.0 * r * vp;
            g_vvr[3 * g + 1] = gl * (-4.0 * r * vs) + gm * 2.0 * r * vs;
            g_vvr[3 * g + 2] = gl * (vp * vp - 2.0 * vs * vs) + gm * vs * vs + gr;
        }
    }
    free(fl); free(fm); free(fr); free(nk);
    ws_free(&w);
    return ierr;
}

/* ------------------------------------------------------------------------ */
/* gnutils.f90:2600-2630 GENHLOC: H += DMAG2 * J^T J per (frequency, source)
 * with J's rows the real parts then the imaginary parts of the saved
 * Jacobian (REAL*4 read into REAL*8), i.e. H(c1, c2) += alpha_j sum_o
 * (Re J(o,c1) Re J(o,c2) + Im J(o,c1) Im J(o,c2)) in FP64. */
void orc_gn_hessian(int njac, const int *nobs, const double *alpha, int nnpgl, const float complex *jac,
                    double *h)
{
    size_t off = 0;
    for (int j = 0; j < njac; j++) {
        const int no = nobs[j];
        const float complex *J = jac + off;
        for (int c2 = 0; c2 < nnpgl; c2++)
            for (int c1 = 0; c1 < nnpgl; c1++) {
                double s = 0.0;
                for (int o = 0; o < no; o++) {
                    float complex a = J[(size_t)c1 * no + o], b = J[(size_t)c2 * no + o];
                    s += (double)crealf(a) * (double)crealf(b) + (double)cimagf(a) * (double)cimagf(b);
                }
                h[(size_t)c2 * nnpgl + c1] += alpha[j] * s;
            }
        off += (size_t)no * nnpgl;
    }
}

/* ------------------------------------------------------------------------ */
/* pchadj.f90:225-281 with NVINV = 1: the loop over local gradient columns,
 * receivers and components (KRHS = IRHS), the COMPLEX*8 inner product of a
 * row of G^T with column INPGL of F (pchadj.f90:239-254), the Jacobian copy
 * (:256-262) and the 1x1 CGEMM('C','N') into HBUFF (:264-281). */
void orc_pchadj_jac(int nnpgl, int nrhs, int lrhs, const int *mygrad, const int *jcsc,
                    const int *icsc, const float complex *g, const float complex *fmat,
                    float complex *jac, float *hdiag)
{
    for (int c = 0; c < nnpgl; c++) {
        int i1 = jcsc[c], i2 = jcsc[c + 1] - 1;
        float complex hsum = 0.0f;
        for (int irhs = 1; irhs <= nrhs; irhs++) {
            float complex acc = 0.0f;
            for (int inz = i1; inz <= i2; inz++) {
                int jdof = icsc[inz - 1];
                if (jdof == 0) continue;
                acc += cmulf_ref(g[(size_t)(irhs - 1) * lrhs + (jdof - 1)], fmat[inz - 1]);
            }
            jac[(size_t)c * nrhs + (irhs - 1)] = acc;
        }
        for (int irhs = 0; irhs < nrhs; irhs++) {
            float complex v = jac[(size_t)c * nrhs + irhs];
            hsum += cmulf_ref(conjf(v), v);
        }
        hdiag[mygrad[c] - 1] += crealf(hsum);
    }
}

/* gnmfree.f90:185-212: JP1(IOBS) = sum_INPGL DCMPLX(J) * DCMPLX(PW(ILOC), 0)
 * (COMPLEX*16), RWORK(ILOC) += DREAL(conj(J) * JP2) and Z2 += ALPHAM * Z1
 * (DAXPY), one Jacobian after the other.  The running sum starts at zero
 * (the reference clears it with DSCAL(0) on fresh memory, gnmfree.f90:60,
 * which is zero unless the heap held NaN there). */
void orc_gnmfree(int njac, const int *nobs, const double *alpha, int nnpgl, const int *mygrad,
                 int na35, const float complex *jac, const double *p, double *z)
{
    size_t off = 0;
    double *rwork = calloc((size_t)na35, sizeof(double));
    int mobs = 0;
    for (int j = 0; j < njac; j++) mobs = nobs[j] > mobs ? nobs[j] : mobs;
    double complex *jp = calloc((size_t)(mobs > 0 ? mobs : 1), sizeof(double complex));
    for (int j = 0; j < njac; j++) {
        const float complex *J = jac + off;
        const int no = nobs[j];
        for (int iobs = 0; iobs < no; iobs++) {
            double complex acc = 0.0;
            for (int c = 0; c < nnpgl; c++) {
                float complex v = J[(size_t)c * no + iobs];
                double pr = p[mygrad[c] - 1];
                acc += CMPLX((double)crealf(v) * pr, (double)cimagf(v) * pr);
            }
            jp[iobs] = acc;
        }
        for (int i = 0; i < na35; i++) rwork[i] = 0.0;
        for (int c = 0; c < nnpgl; c++) {
            int iloc = mygrad[c] - 1;
            for (int iobs = 0; iobs < no; iobs++) {
                float complex v = J[(size_t)c * no + iobs];
                double complex a = jp[iobs];
                rwork[iloc] += (double)crealf(v) * creal(a) + (double)cimagf(v) * cimag(a);
            }
        }
        for (int i = 0; i < na35; i++) z[i] += alpha[j] * rwork[i];
        off += (size_t)no * nnpgl;
    }
    free(rwork);
    free(jp);
}
