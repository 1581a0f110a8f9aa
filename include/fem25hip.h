/*
 * fem25hip.h -- C ABI of the MI355X (gfx950) implementation of fem25's
 * element-operator / assembly / gradient hot path.
 *
 * All arrays are plain host pointers unless an entry point says "_dev".
 * Integer maps keep the reference's conventions: 1-based, column-major
 * (Fortran) layouts, 0 marks a fixed DOF in LM.  Complex arrays are
 * interleaved (re, im) pairs of float (COMPLEX*8) or double (COMPLEX*16).
 * Every entry point returns 0 on success and a non-zero error code
 * otherwise (the reference's IERR convention); none of them aborts.  Calls
 * are synchronous with respect to the caller: results are complete when the
 * call returns.  A mesh handle may be used by one host thread at a time.
 */
#ifndef FEM25HIP_H
#define FEM25HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    F25_OK = 0,
    F25_ERR_ARG = 1,        /* bad argument / size */
    F25_ERR_JACOBIAN = 2,   /* det J <= 2.22e-16 or negative (cjac.f:115-125) */
    F25_ERR_PATTERN = 3,    /* a (row, col) pair is missing from the CSR pattern (asmble.f90:296-300) */
    F25_ERR_HIP = 4,        /* a HIP runtime call failed */
    F25_ERR_NOMEM = 5,
    F25_ERR_GRAPH = 6,      /* numbering failed (graph25.f fidelity checks) */
    F25_ERR_INVTYPE = 7,    /* CINVTYPE other than PP/SS (vforce25.f:66-85) */
    F25_ERR_STATE = 8       /* call order: model or inversion structures missing */
};

/* Precision selector for assembled operators and products. */
enum { F25_C64 = 0, F25_C128 = 1 };
/* Arithmetic of the batched element-block product. */
enum { F25_APPLY_F64 = 0, F25_APPLY_F32 = 1, F25_APPLY_F32_REFINED = 2 };
/* or'ed with F25_APPLY_F64 in f25_apply_batch_dev: the product's launches
 * (one per element colour) are captured once into a HIP graph for the given
 * x, y and npair and the graph is replayed by later calls with the same
 * arguments (omega and ky may change: they are staged before the replay). */
enum { F25_APPLY_GRAPH = 16 };

/*
 * Anchor mesh, the fields of MESH_INFO (fwd_struc.h:10-141) the path reads.
 * pml[4] holds RCOEFF, DMAX, KAPPA, FCENT for body waves (CFTYPE /= 'S');
 * pml_srf[4] the surface-wave set (CFTYPE == 'S'), asmble.f90:78-88.
 */
typedef struct {
    int nord;                /* NORD: GLL order, NLXI = NLETA = NORD+1 */
    int nelem;               /* NELEM */
    int nnpg;                /* NNPG anchors */
    const int *ieng;         /* IENG(4,NELEM), 1-based, CCW */
    const double *xlocs;     /* XLOCS(NNPG) m */
    const double *zlocs;     /* ZLOCS(NNPG) m, z up */
    const double *xd;        /* XD(NNPG) distance into PML, may be NULL without 'A' elements */
    const double *zd;        /* ZD(NNPG) */
    const char *cdomain;     /* CDOMAIN(NELEM): 'I', 'E', 'A' */
    const char *cnnpg;       /* CNNPG(NNPG), CHARACTER(2): 2*NNPG chars */
    double xwidth, zwidth;   /* XWIDTH, ZWIDTH */
    double pml[4];
    double pml_srf[4];
} f25_mesh_desc;

typedef struct f25_mesh_s f25_mesh;
typedef struct f25_graph_s f25_graph;

/* ---- numbering on the host (no device needed) ------------------------- */

/* GEN_GRAPH25's numbering sequence (graph25_driver.f90:30-86): GENIEN
 * (graph25.f:509), GENID (graph25.f:315), GENLM (graph25.f:1), GENCRS_QUICK
 * (graph25.f:136, when with_csr) and CNMESH (bielak_utils.f:934).  sizes:
 * [0] NNP, [1] NDOF, [2] NZ (0 without CSR). */
int f25_graph_build(const f25_mesh_desc *desc, int with_csr, f25_graph **out, int64_t sizes[3]);
/* IEN(NEN,NELEM), LM(3,NEN,NELEM), IRPTR(NDOF+1), JCPTR(NZ), CNP(NDOF); NULL skips. */
int f25_graph_get(const f25_graph *g, int *ien, int *lm, int *irptr, int *jcptr, char *cnp);
int f25_graph_destroy(f25_graph *g);

/* ---- mesh and model -------------------------------------------------- */

/* Builds the GLL numbering with GEN_GRAPH25's semantics
 * (graph25_driver.f90:30-86: GENIEN, GENID, GENLM, CNMESH) and uploads the
 * mesh to the current HIP device.  The CSR pattern (GENCRS_QUICK) is built
 * on first use by the assembled-operator entry points. */
int f25_mesh_create(const f25_mesh_desc *desc, f25_mesh **out);
/* Creates a handle on a numbering the caller already holds: LM(3,NEN,NELEM)
 * and, optionally (irptr == NULL skips), the CSR pattern IRPTR(NDOF+1),
 * JCPTR(NZ).  This is how the drop-in shims hand over MSH%LM, MSH%IRPTR_LOC
 * and MSH%JCPTR_LOC unchanged. */
int f25_mesh_create_with_graph(const f25_mesh_desc *desc, int ndof, const int *lm,
                               const int *irptr, const int *jcptr, f25_mesh **out);
int f25_mesh_destroy(f25_mesh *m);
/* sizes[0] = NNP fine nodes, [1] = NDOF, [2] = NZ (builds the CSR pattern),
 * [3] = number of element colours, [4] = NEN, [5] = quadrature points per
 * direction (2*NORD+5). */
int f25_mesh_sizes(f25_mesh *m, int64_t sizes[6]);
/* Copies the numbering out in the reference's layouts (any pointer may be
 * NULL): IEN(NEN,NELEM), LM(3,NEN,NELEM), IRPTR(NDOF+1), JCPTR(NZ), CNP(NDOF). */
int f25_mesh_graph(f25_mesh *m, int *ien, int *lm, int *irptr, int *jcptr, char *cnp);
/* DENS(NNPG) kg/m^3 and ECOEFF(NNPG,2) = (lambda, mu) Pa.  Recomputes the
 * element blocks on the device; a call with the model already set returns
 * at once (so drop-in shims may call it before every product). */
int f25_set_model(f25_mesh *m, const double *dens, const double *ecoeff);

/* Quality factors QP(NNPG), QS(NNPG) at the anchors and the reference
 * frequency FREQ0 (Hz): the medium becomes dispersive (MESH_INFO's LDISP,
 * asmble.f90:159-168, 187-201) and the assembled operator uses KMDAMP /
 * CMDAMP (estiff25.f:511-961) with VCMPLX velocities (haskattn.f:602-641).
 * qp = NULL returns to the elastic operator.  While set, the batched
 * product integrates every element per omega, and the entry points built
 * on frequency-independent element blocks (gradient, Gauss-Newton
 * products) return F25_ERR_STATE. */
int f25_set_attenuation(f25_mesh *m, const double *qp, const double *qs, double freq0);

/* ---- assembled operator (replaces ASMBLE_DRIVER, asmble.f90:1-231) ------ */

/* S(omega, ky) = K - omega^2 M in the CSR order of the pattern, with
 * omega = 2 pi freq and ky = omega py (estiff25.f:61).  precision F25_C64
 * writes COMPLEX*8 (the reference's CDIST), F25_C128 writes COMPLEX*16.
 * out is a host array of NZ complex values. */
int f25_asmble(f25_mesh *m, char cftype, double freq, double py, int precision, void *out);
/* Same into device memory (NZ complex values), left resident for f25_csr_apply_dev. */
int f25_asmble_dev(f25_mesh *m, char cftype, double freq, double py, int precision, void *out_dev);
/* y = S x with an assembled S (device pointers, NDOF-long complex vectors). */
int f25_csr_apply_dev(f25_mesh *m, int precision, const void *s_dev, const void *x_dev, void *y_dev);

/* Bielak effective force (ASMB_PEFF, asmble.f90:318-377, one rank):
 * PEFF(row 'B') -= S u_E over 'E' columns, PEFF(row 'E') += S u_E over 'B'
 * columns.  cdist, ue, peff are COMPLEX*8 host arrays. */
int f25_asmb_peff(f25_mesh *m, const void *cdist, const void *ue, void *peff);
/* The Bielak sources of a batch, FP64, device memory, [dof][pair] complex128
 * (pair fastest, the batched product's layout):
 *   f25_fill_uwave_dev: FILL_UWAVE_QMESH (disp25.f90:710-908) for nwave
 *     plane waves, u = (cos baz sin aoi, sin baz sin aoi, cos aoi)
 *     exp(-i (kappax x + kappaz z)) at every DOF (angles in degrees, the
 *     DOF placed by the first element holding it, as the reference does);
 *   f25_peff_batch_dev: ASMB_PEFF's effective force for npair (omega, ky)
 *     pairs with ue_dev[:, pair] the wavefield of that pair's source: -S u_E
 *     over 'E' columns on 'B' rows, +S u_E over 'B' columns on 'E' rows, zero
 *     elsewhere, S(omega, ky) from the element blocks (the batched product
 *     on the ring of elements holding both classes).  Not for dispersive
 *     media (F25_ERR_ARG; f25_asmb_peff takes their CDIST per pair). */
int f25_fill_uwave_dev(f25_mesh *m, int nwave, const double *kappax, const double *kappaz, const double *baz,
                       const double *aoi, void *ue_dev);
int f25_peff_batch_dev(f25_mesh *m, int npair, const double *omega, const double *ky, const void *ue_dev,
                       void *peff_dev);
/* Both for a batch of plane-wave sources, one per pair: ue_dev[:, b] the
 * wave (kappax[b], kappaz[b], baz[b], aoi[b]), peff_dev[:, b] its effective
 * force at (omega[b], ky[b]) (the forward right-hand sides of
 * funcgradh.f90:207-233).  flags F25_APPLY_GRAPH: the launches are captured
 * in a HIP graph on the first call and replayed while npair and the two
 * buffers stay the same. */
int f25_bielak_rhs_dev(f25_mesh *m, int npair, const double *omega, const double *ky, const double *kappax,
                       const double *kappaz, const double *baz, const double *aoi, void *ue_dev, void *peff_dev,
                       int flags);

/* ---- batched matrix-free product ---------------------------------------- */

/* y_b = S(omega_b, ky_b) x_b for b = 0..npair-1: the operator ASMBLE_DRIVER
 * assembles (asmble.f90:140-212) with the body-wave PML parameters (CFTYPE
 * other than 'S').  Interior elements run as frequency-independent element
 * blocks (KM25ISOQ, estiff25.f:1-164); PML elements (CM25ISOQ,
 * estiff25.f:168-334) as per-omega element matrices in FP64, rebuilt when the
 * batch's set of omegas changes; in a dispersive medium (f25_set_attenuation)
 * every element runs as per-omega matrices (KMDAMP / CMDAMP,
 * estiff25.f:511-961).  x and y are device arrays of npair*NDOF
 * complex values laid out [dof][pair] (pair fastest), double complex for
 * F25_APPLY_F64 and F25_APPLY_F32_REFINED, float complex for F25_APPLY_F32. */
int f25_apply_batch_dev(f25_mesh *m, int npair, const double *omega, const double *ky,
                        int mode, const void *x_dev, void *y_dev);
/* Residual for iterative refinement of the forward solves (and of the
 * adjoint solves on meshes without PML, where S is Hermitian): the
 * reference factorises S in single precision (CMUMPS, bielak25.f90:203,
 * adjgrad.f90:329); a solution u of that factorisation is refined by
 * u += S32^-1 r with the FP64 residual
 *   r_b = f_b - S(omega_b, ky_b) u_b   (the operator of f25_apply_batch_dev)
 * for every pair, and rnorm[b] = |r_b|_2 / |f_b|_2 (host array of npair).
 * u, f, r: distinct [dof][pair] double complex device arrays. */
int f25_residual_batch_dev(f25_mesh *m, int npair, const double *omega, const double *ky,
                           const void *u_dev, const void *f_dev, void *r_dev, double *rnorm);
/* Host convenience wrapper: x, y are [pair][dof] double complex (one vector
 * per pair, NDOF contiguous, i.e. Fortran X(NDOF,NPAIR)). */
int f25_apply_batch(f25_mesh *m, int npair, const double *omega, const double *ky, int mode,
                    const void *x, void *y);
/* Tuning hook of the FP64 product (no counterpart in the reference, which
 * has no batched product): number of row tiles per output component, 0 =
 * the order's default.  Compiled candidates: p2, p3: 1, 2; p4: 1..3; p5,
 * p6: 2..4.  Rewrites the product's coefficient layout when the model is
 * set.  fem25_amd/tune_apply.py times the candidates. */
int f25_set_apply_tiles(f25_mesh *m, int nrt);

/* ---- gradient ------------------------------------------------------------ */

/* Exact adjoint-state gradient for ntrip (omega, ky) wavefield pairs:
 *   g[3*a + p] += sum_t Re( lam_t^H dS(omega_t, ky_t)/dm_{a,p} u_t )
 * for anchor a and p = (Vp, Vs, rho), model as in f25_set_model
 * (lambda = rho(Vp^2-2Vs^2), mu = rho Vs^2).  Anchors of PML elements hold
 * fixed parameters (the reference inverts none of them, gengmask.f90:110-121):
 * their entries receive nothing; the same holds for f25_gn_fv (their
 * perturbation is ignored), f25_gn_fhw and f25_gn_diag.  u, lam:
 * [trip][dof] double complex host arrays.  With lam solving
 * S^H lam = R^T (d - R u) this is the gradient of 1/2 |R u - d|^2. */
int f25_gradient(f25_mesh *m, int ntrip, const double *omega, const double *ky,
                 const void *u, const void *lam, double *g);
int f25_gradient_dev(f25_mesh *m, int ntrip, const double *omega, const double *ky,
                     const void *u_dev, const void *lam_dev, double *g_dev);

/* The same gradient from complex64 fields (u, lam: [trip][dof] float complex
 * device arrays): the interpolation to the quadrature points and the
 * densities of each triple in FP32, their sums over triples, points,
 * elements and the chain rule in FP64; g FP64 as above. */
int f25_gradient32_dev(f25_mesh *m, int ntrip, const double *omega, const double *ky,
                       const void *u_dev, const void *lam_dev, double *g_dev);

/* ---- reference virtual-force path (CFMAT_DIST / ADJFLAM_DIST) ------------ */

/* GENGMASK + ELEM_WTS + GRADPTRS for one rank (gengmask.f90:1-262,
 * 758-841, 962-1198), no mask file.  cinvtype "PP" or "SS" (NVINV = 1) or
 * "PS" (NVINV = 2: two columns per anchor, P then S, with the same rows and
 * MYGRAD = 2 (INPINV - 1) + IVINV, gn25.f90:170-183); mrdof1 holds
 * MRDOF(1,1:NREC) (used only when linrec == 0).  ICSC_FDIST is sized
 * exactly: the reference's GRADPTRS bounds it by NNPG*NCON*NEE and stops
 * ('ICSC_FDIST too small', gengmask.f90:1069-1073) when a PS inversion
 * covers more than about 60% of the anchors.  sizes: [0] NNPINV,
 * [1] NCON, [2] NNPGL, [3] NZ_FDIST, [4] MBUFRHS, [5] NA35. */
int f25_inv_setup(f25_mesh *m, const char *cinvtype, double vpvs, int linrec, int lkbdry,
                  int nrec, const int *mrdof1, int sizes[6]);
/* MASKG(NNPG), MYGRAD(NNPGL), JCSC_FDIST(NNPGL+1), ICSC_FDIST(NZ_FDIST),
 * ELEM_WTS(NELEM); any pointer may be NULL. */
int f25_inv_get(f25_mesh *m, int *maskg, int *mygrad, int *jcsc, int *icsc, double *elem_wts);
/* Inversion structures as the caller already holds them (the INV_INFO
 * fields fwd_struc.h:249-368 after GENGMASK/GRADPTRS/ELEM_WTS, whatever mask
 * file built them), for the drop-in CFMAT_DIST: MCONN(NNPG,NCON),
 * MASKG(NNPG), MYGRAD(NNPGL), JCSC_FDIST(NNPGL+1), ICSC_FDIST(NZ_FDIST),
 * ELEM_WTS(NELEM), all 1-based.  Column INPGL inverts at the anchor INPG
 * with (MASKG(INPG) - 1) NVINV + IVINV == MYGRAD(INPGL) (hessloc.f90:66-72).
 * NVINV is 1 for CINVTYPE PP or SS and 2 for PS (F25_ERR_INVTYPE else). */
int f25_inv_attach(f25_mesh *m, const char *cinvtype, double vpvs, int nvinv, int ncon,
                   const int *mconn, const int *maskg, int nnpgl, const int *mygrad, const int *jcsc,
                   int nz_fdist, const int *icsc, int na35, const double *elem_wts);
/* CFMAT_DIST (hessloc.f90:1-133 + vforce25.f:1-369): FMAT_DIST(NZ_FDIST)
 * COMPLEX*8 from WAVE(NDOF) COMPLEX*8, reference semantics.  PS returns
 * F25_ERR_INVTYPE, as VFORCE25 does (vforce25.f:81-85: no joint P and S
 * derivatives). */
int f25_cfmat_dist(f25_mesh *m, double freq, double py, const void *wave, void *fmat);
/* ADJFLAM_DIST (adjgrad.f90:88-145, one rank): GPROD(NA35) REAL*4. */
int f25_adjflam_dist(f25_mesh *m, const void *fmat, const void *adjlam, float *gprod);

/* ---- handle-free sparse forms -------------------------------------------- */

/* ASMB_PEFF's local product (asmble.f90:348-374) on the caller's sparse
 * structures, before the reduction: BUFF(NDOF) = 0, then for local row
 * IDOFL (global DOF MYDOFS(IDOFL)) with CNP 'B'/'E', BUFF(IDOF) -= S u_E over
 * 'E' columns of a 'B' row and += S u_E over 'B' columns of an 'E' row.
 * CNP(NDOF) chars, MYDOFS(NDOFL), IRPTR_LOC(NDOFL+1), JCPTR_LOC(NZLOC) 1-based;
 * cdist COMPLEX*8(NZLOC), ue and buff COMPLEX*8(NDOF) host arrays.  The caller
 * sums BUFF over ranks (the reference's MPI_REDUCE, asmble.f90:374). */
int f25_asmb_peff_csr(int ndof, int ndofl, int nzloc, const char *cnp, const int *mydofs,
                      const int *irptr_loc, const int *jcptr_loc, const void *cdist, const void *ue,
                      void *buff);
/* ADJFLAM_DIST's local product (adjgrad.f90:127-139) before the reduction:
 * VWORK(NA35) = 0, VWORK(MYGRAD(c)) -= sum over column c of
 * Re(conj(FMAT_DIST(i)) ADJLAM(ICSC_FDIST(i))).  fmat COMPLEX*8(NZ_FDIST),
 * adjlam COMPLEX*8(NDOF), vwork REAL*4(NA35) host arrays; MYGRAD entries
 * distinct.  The caller reduces VWORK onto its master (adjgrad.f90:141). */
int f25_adjflam_csc(int nnpgl, int ndof, int nz_fdist, int na35, const int *mygrad, const int *jcsc,
                    const int *icsc, const void *fmat, const void *adjlam, float *vwork);

/* ---- Jacobians and the Gauss-Newton contraction --------------------------- */

/* PCHADJ's element-sparse product for one (frequency, source)
 * (pchadj.f90:225-281, NVINV = 1) on the handle's inversion structures:
 *   JACL(irhs, c) = sum over column c of F of
 *                   GRNSFNS((irhs-1)*LRHS + ICSC_FDIST(i)) * FMAT_DIST(i)
 * for irhs = 1..NRHS (receivers x components), written to jac as
 * JACL(NRHS, NNPGL) COMPLEX*8 column-major (RCJLOC + i CCJLOC), and
 * hdiag(MYGRAD(c)) += sum_irhs |JACL(irhs, c)|^2 (REAL*4, NA35: the 1x1
 * CGEMM('C','N') of the gradient pre-conditioner).  grnsfns COMPLEX*8
 * (LRHS*NRHS, LRHS >= NDOF), fmat COMPLEX*8 (NZ_FDIST); host arrays. */
int f25_pchadj_jac(f25_mesh *m, int nrhs, int lrhs, const void *grnsfns, const void *fmat, void *jac,
                   float *hdiag);
/* Same on device arrays; gt_dev holds G transposed: NDOF rows of NRHS
 * COMPLEX*8 ([dof][rhs]). */
int f25_pchadj_jac_dev(f25_mesh *m, int nrhs, const void *gt_dev, const void *fmat_dev, void *jac_dev,
                       float *hdiag_dev);
/* GNMFREE's contraction (gnmfree.f90:185-212) for a batch of njac
 * Jacobians with uneven observation counts, packed one after the other
 * (J_j: nobs[j] x NNPGL COMPLEX*8, column-major, as PCHADJ writes them):
 *   z(MYGRAD(c)) += sum_j alpha[j] Re( J_j(:,c)^H (J_j p(MYGRAD)) )
 * in FP64 (alpha = 1, or |source|^2 when migrating).  z and p hold NA35
 * values.  nobs and alpha are host arrays in both forms. */
int f25_gnmfree(f25_mesh *m, int njac, const int *nobs, const double *alpha, const void *jac,
                const double *p, double *z);
int f25_gnmfree_dev(f25_mesh *m, int njac, const int *nobs, const double *alpha, const void *jac_dev,
                    const double *p_dev, double *z_dev);

/* The Gauss-Newton Hessian over the local gradient columns, the block
 * GENHLOC accumulates from the saved Jacobians (gnutils.f90:2408-2660:
 * J^T J with DMAG2 = |STF|^2, the Jacobian's real and imaginary parts as
 * separate rows, pchadj.f90:256-262):
 *   H(c1, c2) += sum_j alpha[j] Re( J_j(:, c1)^H J_j(:, c2) ),  c1, c2 = 1..NNPGL
 * H: NNPGL x NNPGL FP64 column-major device array, symmetric on entry
 * (e.g. zero) and on return.  jac_dev packed as for f25_gnmfree (COMPLEX*8
 * on the device); nobs, alpha (>= 0) host arrays.  FP64 products of the
 * FP32 entries (exact) summed by FP64 MFMA (rocBLAS DSYRK). */
int f25_gn_hessian_dev(f25_mesh *m, int njac, const int *nobs, const double *alpha, const void *jac_dev,
                       double *h_dev);
/* The Jacobian file of one (frequency, source) in the record layout PCHADJ
 * writes (pchadj.f90:371-395, Fortran sequential unformatted: 4-byte record
 * lengths): (NRHSF, NNPGL, NDIM, NREC), LOBS(NDIM, NREC) LOGICAL*4 as given,
 * Re J and Im J (NRHSF x NNPGL REAL*4, rows fastest), MYGRAD(NNPGL).  jac:
 * JACL(NRHSF, NNPGL) COMPLEX*8 as f25_pchadj_jac writes it, on the device
 * when jac_on_device != 0.  path: the caller's file name (the reference uses
 * ./scratch/jac_<rank>-<FREQ F12.5>-<ISRC>.dat). */
int f25_write_jac(const char *path, int nrhsf, int nnpgl, int ndim, int nrec, const int *lobs, const void *jac,
                  int jac_on_device, const int *mygrad);
/* f25_gnmfree_dev without any allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * per-call tables, the J p vectors and their column-chunk partial sums live
 * in a 32 MiB device array of the library, one per GPU, held until the
 * process ends; the call returns F25_ERR_ARG when
 *   sum(nobs) * (min(256, ceil(NNPGL / 64)) + 1) * 16 bytes + 32 * njac
 * exceeds it (split the Jacobians over calls: z accumulates).  Results are
 * bit for bit those of f25_gnmfree_dev.  Calls on one GPU share the array:
 * calls that overlap on one GPU (other host threads, other handles) race. */
int f25_gnmfree_dev_noalloc(f25_mesh *m, int njac, const int *nobs, const double *alpha, const void *jac_dev,
                            const double *p_dev, double *z_dev);
/* GNMFREE's two products on caller-held arrays (no handle), for a caller
 * that sums JP over ranks in between as gnmfree.f90:195-196 does:
 *   f25_gn_jp:  JP_j(obs) = sum_c J_j(obs,c) p(MYGRAD(c)), jp COMPLEX*16,
 *               the njac vectors packed (sum of nobs values);
 *   f25_gn_jhz: z(MYGRAD(c)) += sum_j alpha[j] Re( J_j(:,c)^H JP_j ).
 * MYGRAD(NNPGL) 1-based and distinct, p and z REAL*8 (NA35), jac packed as
 * for f25_gnmfree; host arrays. */
int f25_gn_jp(int nnpgl, int na35, const int *mygrad, int njac, const int *nobs, const void *jac,
              const double *p, void *jp);
int f25_gn_jhz(int nnpgl, int na35, const int *mygrad, int njac, const int *nobs, const double *alpha,
               const void *jac, const void *jp, double *z);

/* ---- adjoint-source pipeline (between the forward and the adjoint solve) ---- */

/* EXRESP25D (locrec2d.f:867-925): est(i, irec, isrc) = recv(i, irec) *
 * sol((isrc-1)*ndof + mrdof(i, irec)) * exp(+i 2 pi freq py(isrc) yrec(irec)),
 * zero where mrdof is 0.  sol COMPLEX*8 (ndof*nsrc), mrdof (3, nrec) in
 * 0..ndof, recv (3, nrec) COMPLEX*8, py (nsrc), yrec (nrec), est (3, nrec,
 * nsrc) COMPLEX*8.  Host arrays; the _dev form takes sol and est on the
 * device (the receiver tables stay host arrays). */
int f25_exresp25d(int ndof, int nsrc, int nrec, double freq, const double *py, const double *yrec,
                  const void *sol, const int *mrdof, const void *recv, void *est);
int f25_exresp25d_dev(int ndof, int nsrc, int nrec, double freq, const double *py, const double *yrec,
                      const void *sol_dev, const int *mrdof, const void *recv, void *est_dev);
/* CADJLAM's right-hand side (adjgrad.f90:229-300, without the data
 * covariance option LCOVD, for which see f25_cadjlam_rhs_covd) for npair
 * (ifreq, isrc) pairs (1-based, with their frequency in Hz and slowness
 * py): estimates rotated to (N,E,Z) by
 * azmod degrees (magnitude/phase pairs turned complex first when lunwrap),
 * data and estimates shifted by exp(-i 2 pi freq py yrec), BPRESID4
 * residuals (objfn.f:757-839: norm, irestp, epss, dtmax), rotated back,
 * RESID_DRHS (adjgrad.f90:448-523: HDIFFER(ibphase), scatter at mrdof,
 * conjugate).  est, obs COMPLEX*8 and wghts REAL*4 in the reference's
 * (3, NFREQ, NREC, NSRC) layout; rhs (ndof, npair) COMPLEX*8, the dense
 * right-hand side the transpose solve takes.  iwarn = 1 when a receiver DOF
 * lies outside 1..ndof (skipped, as the reference does). */
int f25_cadjlam_rhs(int ndof, int nfreq, int nrec, int nsrc, int npair, const int *ifreq, const int *isrc,
                    const double *freq, const double *py, double azmod, int norm, int irestp, float epss,
                    float dtmax, int ibphase, int lunwrap, const double *yrec, const int *mrdof, const void *est,
                    const void *obs, const float *wghts, void *rhs, int *iwarn);
int f25_cadjlam_rhs_dev(int ndof, int nfreq, int nrec, int nsrc, int npair, const int *ifreq, const int *isrc,
                        const double *freq, const double *py, double azmod, int norm, int irestp, float epss,
                        float dtmax, int ibphase, int lunwrap, const double *yrec, const int *mrdof,
                        const void *est_dev, const void *obs_dev, const float *wghts_dev, void *rhs_dev,
                        int *iwarn);

/* CADJLAM's right-hand side with the data covariance (LCOVD,
 * adjgrad.f90:219-235 and 296-298): per (ifreq, isrc) pair the roughening
 * matrix of LAPLACE_1DV2 (locrec2d.f:334-509; with ldwght the element
 * lengths are the summed normalised spacings dx between observed receivers,
 * else 1) and PERM_OBS's ordering (locrec2d.f:511-549) of the untransformed
 * observations, then BPRESID_COVD4 (objfn.f:635-733: CRESID(irestp) with
 * the single-precision phase window dtmax, weighted on both sides by the
 * receiver weights) in place of BPRESID4; the rotations, phase shift and
 * RESID_DRHS as in f25_cadjlam_rhs.  dx holds the nrec - 1 spacings
 * (f25_rec_space1d), each positive and finite when ldwght (F25_ERR_ARG
 * otherwise; the reference stops on a zero spacing sum).  The roughening
 * matrix is tridiagonal and is applied as such. */
int f25_cadjlam_rhs_covd(int ndof, int nfreq, int nrec, int nsrc, int npair, const int *ifreq, const int *isrc,
                         const double *freq, const double *py, double azmod, int irestp, float dtmax, int ibphase,
                         int lunwrap, int ldwght, const double *dx, const double *yrec, const int *mrdof,
                         const void *est, const void *obs, const float *wghts, void *rhs, int *iwarn);
int f25_cadjlam_rhs_covd_dev(int ndof, int nfreq, int nrec, int nsrc, int npair, const int *ifreq,
                             const int *isrc, const double *freq, const double *py, double azmod, int irestp,
                             float dtmax, int ibphase, int lunwrap, int ldwght, const double *dx,
                             const double *yrec, const int *mrdof, const void *est_dev, const void *obs_dev,
                             const float *wghts_dev, void *rhs_dev, int *iwarn);
/* REC_SPACE1D (locrec2d.f:157-223): dx[i] = |xrec[i+1] - xrec[i]| divided by
 * the smallest such spacing.  F25_ERR_ARG, dx zeroed, for one receiver, a
 * spacing below 1e-10 or receivers out of order (the reference's checks). */
int f25_rec_space1d(int nrec, const double *xrec, double *dx);

/* ---- second-order products of QPMFREE (truncated-Newton Hessian) ----------- */

/* CADJ_NU_ALPHA's element loops (qpmfree.f90:622-912) on the attached
 * inversion structures (f25_inv_setup / f25_inv_attach; PP, SS or the joint
 * PS with both of an anchor's variables, qpmfree.f90:677-693, 836-851) and
 * the model: with M_{c,e} = SNGL(ELEM_WTS(e)) V(MYGRAD(c)) dS_e/dm_c, the
 * DABMKISOB25 matrix (dstiff25.f:1-203, P or S form by column c's
 * variable) of element e collocated on column c's anchor,
 *   rhsbnu = sum_c sum_{e in MCONN} M_{c,e}^H adjlam     (RHSBNU, :868-869)
 *   rhsalp = sum_c sum_{e in MCONN} M_{c,e} wave         (RHSALP, :870-871)
 *   p3(MYGRAD(c)) = -Re sum_rows conj(M_{c,last} wave) adjlam   (P3BUF)
 * before the MUMPS solves and MPI reductions the caller keeps (:914-980).
 * v REAL*4 (NA35); adjlam, wave, rhsbnu, rhsalp COMPLEX*8 (NDOF); p3 REAL*4
 * (NA35, zero outside the local columns).  Host arrays; FP64 inside. */
int f25_qp_nu_alpha(f25_mesh *m, double freq, double py, const float *v, const void *adjlam,
                    const void *wave, void *rhsbnu, void *rhsalp, float *p3);
/* MET225_MOD's element loops (qpmfree.f90:1000-1215; PP, SS or PS as
 * f25_qp_nu_alpha): with T1, T2 the last
 * connected element's SNGL(ELEM_WTS(e)) dS_e/dm_c applied to wave and alpha,
 *   p1(MYGRAD(c)) = -Re sum_rows conj(T1) nu,  p2 = -Re sum_rows conj(T2) adjlam
 * (PBUF1, PBUF2 before the reductions).  COMPLEX*8 vectors (NDOF), REAL*4
 * outputs (NA35). */
int f25_met225(f25_mesh *m, double freq, double py, const void *adjlam, const void *nu,
               const void *alpha, const void *wave, float *p1, float *p2);

/* ---- Gauss-Newton element products ---------------------------------------- */

/* F v = dS/dm[v] u for a model perturbation v (NNPG,3) = (dVp, dVs, drho)
 * at anchors, ntrip (omega, ky, u) triples.  u and out: [trip][dof] double
 * complex host arrays. */
int f25_gn_fv(f25_mesh *m, int ntrip, const double *omega, const double *ky, const void *u,
              const double *v, void *out);
/* Re(F^H w): the gradient form with lam := w; g (NNPG,3) accumulated. */
int f25_gn_fhw(f25_mesh *m, int ntrip, const double *omega, const double *ky, const void *u,
               const void *w, double *g);
/* Pseudo-Hessian diagonal: diag[a,p] += sum_t || dS_t/dm_{a,p} u_t ||^2. */
int f25_gn_diag(f25_mesh *m, int ntrip, const double *omega, const double *ky, const void *u,
                double *diag);

/* ---- diagnostics ------------------------------------------------------------ */
const char *f25_last_error(void);
/* Runs one 16x16x4 MFMA in FP64 and FP32 on the current device and checks
 * the operand/result lane layout the batched product relies on; *maxerr is
 * the largest deviation (0 when the layout holds). */
int f25_selftest_mfma(double *maxerr);
/* Workgroup residency on the current device: *groups_per_cu = the most
 * workgroups of nt threads (192, 256, 320, 384 or 512) holding vgprs
 * registers per lane (128 or 168) and lds bytes of LDS that are resident
 * at once on one CU (median over the CUs), measured with spinning probe
 * workgroups.  The launch shapes of the gradient, element-derivative and
 * product kernels are chosen for what it reports (a workgroup of four or
 * more waves holds ceil(waves / 4) wave slots on every SIMD of its CU).
 * No counterpart in the reference (a diagnostic of this library). */
int f25_selftest_residency(int nt, int vgprs, int lds, int *groups_per_cu);
/* Duration (ms) of the kernels of the most recent call of the named family
 * ("apply", "grad", "asmble", "cfmat", "gn_diag", "pchadj", "gnmfree"),
 * measured with HIP events on the stream the kernels ran on; -1 if
 * unknown. */
double f25_last_kernel_ms(const char *family);

#ifdef __cplusplus
}
#endif
#endif
