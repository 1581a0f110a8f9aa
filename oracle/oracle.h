/*
 * oracle.h -- CPU restatement of fem25's element/assembly/gradient hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * path in fem25_amd/.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product never links or calls it.
 *
 * Conventions follow the reference (Fortran, column-major, 1-based values in
 * integer maps):
 *   IENG(4,NELEM)          ieng[k + 4*e]              anchor numbers, 1-based
 *   LM(3,NEN,NELEM)        lm[i + 3*(a + NEN*e)]      DOF numbers, 1-based, 0 = fixed
 *   IRPTR(NDOF+1), JCPTR   1-based CSR, columns sorted ascending per row
 *   ECOEFF(NNPG,2)         ecoeff[g + NNPG*c]         c=0 lambda, c=1 mu
 *   SHL/SHG(3,NEN,Q1,Q1)   shl[c + 3*(a + NEN*(ix + Q1*iz))]
 *   ME/KE(NEE,NEE)         m[ip + NEE*iq]
 *   CDOMAIN                one char per element ('I','E','A')
 *   CNNPG                  two chars per anchor ("II","FS","FL",...)
 *
 * Parity status of each function is recorded in oracle.c's header.
 */
#ifndef FEM25_ORACLE_H
#define FEM25_ORACLE_H
#include <complex.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* gaussq.f:5 (kind=5, alpha=beta=0, kpts=2): Gauss-Lobatto-Legendre nodes, weights */
void orc_gll(int n, double *x, double *w);
/* dshl.f:1-243: SHL(3,NEN,NINT,NINT) on GLL interpolation points, GLL quadrature */
void orc_dshl(int nl, int nint, double *xipts, double *xigll, double *wgll, double *shl);
/* csf2d.f:1-131 (NGNOD=4) */
int orc_csf2d(double xi, double eta, double *sf, double *dsf /* dsf[2*k+d] */);
/* cjac.f:1-138 CJMK_TVI (isotropic) */
int orc_cjmk_tvi(int nen, int nint, const int *ieng_e, const double *xigll,
                 const double *xlocs, const double *zlocs, const double *dens,
                 const double *ecoeff, int nnpg, const double *shl,
                 double *det, double *rho, double *dmat /* (nint,nint,2) */, double *shg);
/* cjac.f:818-921 ZDAMPFN */
double complex orc_zdampfn(double rn, double vp, double delta, double dmax_in,
                           double rcoeff, double rkappa, double fcent, double omega);
/* cjac.f:399-581 CJACCQ */
int orc_cjaccq(int nen, int nint, double omega, double xwidth, double zwidth,
               double rcoeff, double dmax_in, double rkappa, double fcent,
               const int *ieng_e, const double *xigll, const double *xd, const double *zd,
               const double *xlocs, const double *zlocs, const double *dens,
               const double *ecoeff, int nnpg, const double *shl,
               double *det, double *rho, double *dmat, double *shg,
               double complex *dampx, double complex *dampz);
/* estiff25.f:1-164 KM25ISOQ */
void orc_km25isoq(int nen, int nint, double omega, double py, const double *wts,
                  const double *vlam, const double *vmu, const double *rho,
                  const double *det, const double *shg, double *me, double complex *ke);
/* estiff25.f:168-334 CM25ISOQ */
void orc_cm25isoq(int nen, int nint, double omega, double py, const double *wts,
                  const double *vlam, const double *vmu, const double *rho,
                  const double *det, const double *shg, const double complex *dampx,
                  const double complex *dampz, double *me, double complex *ce);

/* ---- graph (graph25.f) ---- */
typedef struct {
    int nord, nelem, nnpg;
    const int *ieng;            /* (4,nelem) 1-based */
    const double *xlocs, *zlocs;
    const char *cdomain;        /* nelem chars */
    const char *cnnpg;          /* 2*nnpg chars */
} orc_mesh;

/* graph25.f:509-698 GENIEN -> ien(NEN,NELEM) 1-based, returns nnp (<0 on error) */
int orc_genien(const orc_mesh *m, const double *xipts, int *ien);
/* graph25.f:315-505 GENID -> id(3,nnp) 1-based (0 fixed); returns ndof (<0 error) */
int orc_genid(const orc_mesh *m, const double *xipts, const int *ien, int nnp, int *id);
/* graph25.f:1-37 GENLM */
void orc_genlm(int nen, int nelem, const int *id, const int *ien, int *lm);
/* graph25.f:41-86 CONNECT + ICMCON graph25.f:90-131; returns ncon, mconn(nnp,ncon) */
int orc_icmcon(int ngnod, int nnp, int nelem, const int *ien);
void orc_connect(int ngnod, int nnp, int nelem, int ncon, const int *ien, int *mconn);
/* graph25.f:136-266 GENCRS_QUICK; returns nzero; irptr(ndof+1), jcptr(cap) */
long orc_gencrs(int nen, int nnp, int nelem, int ndof, int ncon, const int *mconn,
                const int *ien, const int *lm, int *irptr, int *jcptr, long cap);
/* bielak_utils.f:934-1025 CNMESH */
int orc_cnmesh(int nen, int ndof, int nelem, const char *cdomain, const int *lm, char *cnp);

/* ---- assembly (asmble.f90) ---- */
typedef struct {
    int nord, nelem, nnpg, ndof;
    const int *ieng, *lm, *irptr, *jcptr;
    const double *xlocs, *zlocs, *dens, *ecoeff, *xd, *zd;
    const char *cdomain;
    double xwidth, zwidth;
    double rcoeff, dmax, kappa, fcent;   /* already selected by CFTYPE */
    /* dispersive medium (LDISP): QP, QS at anchors and FREQ0 (mesh_utils) */
    int ldisp;
    const double *qp, *qs;
    double freq0;
} orc_asm;

/* haskattn.f:602-641 VCMPLX */
double complex orc_vcmplx(double freq, double freq0, double v, double q);
/* cjac.f:142-250 CJKM_QUAL; with xd, zd (and dampx, dampz) non-NULL the
 * CJCM_QUAL variant (cjac.f:585-720) for PML elements */
int orc_cjkm_qual(int nen, int nint, const int *ieng_e, const double *xigll,
                  const double *xlocs, const double *zlocs, const double *dens,
                  const double *qp, const double *qs, const double *ecoeff, int nnpg,
                  const double *shl, double omega, double xwidth, double zwidth, double rcoeff,
                  double dmax_in, double rkappa, double fcent, const double *xd, const double *zd,
                  double *det, double *vp, double *vs, double *qpp, double *qss, double *shg,
                  double complex *dampx, double complex *dampz);
/* estiff25.f:511-752 KMDAMP (dampx = dampz = NULL) / 756-961 CMDAMP */
void orc_kmdamp(int nen, int nint, double freq, double freq0, double py, const double *wts,
                const double *vp, const double *vs, const double *qpp, const double *qss,
                const double *det, const double *shg, const double complex *dampx,
                const double complex *dampz, double *me, double complex *ce);
/* asmble.f90:1-231 + asmble.f90:235-314: COMPLEX*8 CDIST in CSR order */
int orc_asmble_cdist(const orc_asm *a, double freq, double py, float complex *cdist);
/* same element loop, FP64 accumulation (disp25.f90:912-989 ASM_DISP style, ME+KE) */
int orc_asmble_z(const orc_asm *a, double freq, double py, double complex *s);
/* y = S x, CSR (loop shape asmble.f90:357-373 without the CNP filter) */
void orc_spmv_z(int ndof, const int *irptr, const int *jcptr, const double complex *s,
                const double complex *x, double complex *y);
void orc_spmv_c(int ndof, const int *irptr, const int *jcptr, const float complex *s,
                const float complex *x, float complex *y);
/* asmble.f90:318-377 ASMB_PEFF (single rank: no reduce) */
void orc_asmb_peff(int ndof, const char *cnp, const int *irptr, const int *jcptr,
                   const float complex *cdist, const float complex *ue, float complex *peff);

/* ---- gradient (djac.f, dstiff25.f, vforce25.f, hessloc.f90, adjgrad.f90, gengmask.f90) ---- */
int orc_djacab(int nen, int nint, int inpg, const int *ieng_e, const double *xigll,
               const double *xlocs, const double *zlocs, const double *dens,
               const double *ecoeff, int nnpg, const double *shl,
               double *det, double *dab /* (nint,nint,2) */, double *shg);
void orc_dkm_vpvs(int nen, int nint, double omega, double py, const double *wts,
                  const double *dvp, const double *dvs, const double *det, const double *shg,
                  double *dme, double complex *dke);

typedef struct {
    int nnpinv, na35, nvinv, ncon, nnpgl, nz_fdist, mbufrhs;
    int *mconn;       /* (nnpg,ncon) anchor connectivity */
    int *maskg;       /* nnpg */
    int *mygrad;      /* nnpgl */
    int *jcsc;        /* nnpgl+1 */
    int *icsc;        /* nz_fdist */
    double *elem_wts; /* nelem */
} orc_inv;

/* gengmask.f90:1-262 GENGMASK with no .mask file, single partition, LINREC/LKBDRY given */
int orc_gengmask(const orc_asm *a, const char *cnnpg, int nvinv, int linrec, int lkbdry,
                 int nrec, const int *mrdof1, orc_inv *inv);
/* gengmask.f90:962-1198 GRADPTRS (single rank) */
int orc_gradptrs(const orc_asm *a, orc_inv *inv);
/* gengmask.f90:758-841 ELEM_WTS */
int orc_elem_wts(const orc_asm *a, orc_inv *inv);
void orc_inv_free(orc_inv *inv);
/* hessloc.f90:1-133 CFMAT_DIST + vforce25.f:1-369 VFORCE25 */
int orc_cfmat_dist(const orc_asm *a, const orc_inv *inv, const char *cinvtype, double vpvs,
                   double freq, double py, const float complex *wave, float complex *fmat);
/* adjgrad.f90:88-145 ADJFLAM_DIST (single rank) */
void orc_adjflam_dist(int nnpgl, int na35, const int *mygrad, const int *jcsc, const int *icsc,
                      const float complex *fmat, const float complex *adjlam, float *gprod);
int orc_djacab2(int nen, int nint, int inpg, const int *ieng_e, const double *xigll,
                const double *xlocs, const double *zlocs, const double *dens,
                const double *ecoeff, int nnpg, const double *shl,
                double *det, double *dab, double *dab2, double *shg);
void orc_dabmkisob25(int nen, int nint, int lisp, int liss, double omega, double py,
                     const double *wts, const double *det, const double *dab, const double *shg,
                     double *dmda, double *dmdb, double complex *dkda, double complex *dkdb);
void orc_dabm2kisob25(int nen, int nint, int lisp, int liss, double omega, double py,
                      const double *wts, const double *det, const double *dab, const double *dab2,
                      const double *shg, double *dmda, double *dmdb, double complex *dkda,
                      double complex *dkdb, double *dmda2, double *dmdb2, double complex *dkda2,
                      double complex *dkdb2);
int orc_qp_nu_alpha(const orc_asm *a, const orc_inv *inv, const char *cinvtype, double freq,
                    double py, const float *v, const float complex *adjlam,
                    const float complex *wave, float complex *rhsbnu, float complex *rhsalp,
                    float *p3);
int orc_met225(const orc_asm *a, const orc_inv *inv, const char *cinvtype, double freq, double py,
               const float complex *adjlam, const float complex *nu, const float complex *alpha,
               const float complex *wave, float *p1, float *p2);
void orc_exresp25d(int ndof, int nsrc, int nrec, double freq, const double *py, const double *yrec,
                   const float complex *sol, const int *mrdof, const float complex *recv,
                   float complex *est);
void orc_crotate(float theta, float complex q1, float complex q2, float complex *rq1, float complex *rq2);
float complex orc_cphm2cm(float xmag, float phase);
float complex orc_cresid(int irestp, float complex obs, float complex est);
void orc_bpresid4(int nrec, int norm, int irestp, float epss, double freq, float dtmax,
                  const float *swght, const float complex *obs, const float complex *est,
                  float complex *resid);
float complex orc_hdiffer(int ibphase, double freq);
void orc_resid_drhs(int ndof, int nrec, int ibphase, double freq, const int *mrdof,
                    const float complex *resid, float complex *rhs, int *iwarn);
void orc_cresid_n(int n, int irestp, const float complex *obs, const float complex *est, float complex *out);
void orc_cphm2cm_n(int n, const float *xmag, const float *ph, float complex *out);
void orc_crotate_n(int n, float theta, const float complex *q1, const float complex *q2, float complex *r1,
                   float complex *r2);
void orc_hdiffer_v(int ibphase, double freq, float *out);
void orc_cadjlam_rhs(int ndof, int nfreq, int nrec, int ifreq, int isrc, double freq, double py,
                     double azmod, int norm, int irestp, float epss, float dtmax, int ibphase,
                     int lunwrap, const double *yrec, const int *mrdof, const float complex *est,
                     const float complex *obs, const float *wghts, float complex *rhs, int *iwarn);
int orc_fill_uwave(const orc_asm *a, double kappax, double kappaz, double baz, double aoi,
                   double complex *ue);
int orc_rec_space1d(int nrec, const double *xrec, double *dx);
int orc_laplace_1dv2(int nrec, int ldwght, const double *dx, const float complex *obs, double *covd);
int orc_perm_obs(int nrec, const float complex *obs, int *iperm);
void orc_bpresid_covd4(int nrec, const int *iperm, int irestp, float dtmax, double freq,
                       const double *covd8, const float *swght, const float complex *obs,
                       const float complex *est, float complex *resid);
int orc_cadjlam_rhs_covd(int ndof, int nfreq, int nrec, int ifreq, int isrc, double freq, double py,
                         double azmod, int irestp, float dtmax, int ibphase, int lunwrap, int ldwght,
                         const double *dx, const double *yrec, const int *mrdof, const float complex *est,
                         const float complex *obs, const float *wghts, float complex *rhs, int *iwarn);

/* pchadj.f90:225-281 (NVINV = 1), one (frequency, source):
 *   JACL(irhs, c) = sum over column c of F of G((irhs-1)*lrhs + ICSC(i)) FMAT(i)
 * in COMPLEX*8 arithmetic; jac[c*nrhs + irhs] (RCJLOC/CCJLOC(KRHS, INPGL));
 * hdiag(MYGRAD(c)) += REAL(sum_irhs conj(JACL) JACL) (the 1x1 CGEMM). */
void orc_pchadj_jac(int nnpgl, int nrhs, int lrhs, const int *mygrad, const int *jcsc,
                    const int *icsc, const float complex *g, const float complex *fmat,
                    float complex *jac, float *hdiag);
/* gnmfree.f90:95-212, one rank and one frequency group: for Jacobians j =
 * 0..njac-1 packed one after the other (nobs[j] x nnpgl, column-major),
 *   JP = J p(MYGRAD) (COMPLEX*16), RWORK(MYGRAD(c)) = Re(J(:,c)^H JP),
 *   z += alpha[j] RWORK. */
/* GENHLOC's local J^T J (gnutils.f90:2600-2630) on packed COMPLEX*8
 * Jacobians: h (nnpgl x nnpgl, column-major) += sum_j alpha_j Re(J_j^H J_j) */
void orc_gn_hessian(int njac, const int *nobs, const double *alpha, int nnpgl, const float complex *jac,
                    double *h);
void orc_gnmfree(int njac, const int *nobs, const double *alpha, int nnpgl, const int *mygrad,
                 int na35, const float complex *jac, const double *p, double *z);

/* exact derivative of the assembled operator with respect to anchor
 * (lambda, mu, rho) and the chain rule to (Vp, Vs, rho):
 *   g[3*(anchor) + {0,1,2}] += Re( lam^H  dS/dm  u )  for interior (non-PML) elements */
int orc_grad_exact(const orc_asm *a, double freq, double py, const double complex *u,
                   const double complex *lam, double *g_lmr, double *g_vvr);

#ifdef __cplusplus
}
#endif
#endif
