// C ABI (include/fem25hip.h): mesh handles, device residency and launches.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "f25_internal.h"
#include "kernels.hip"

using namespace f25;
using namespace f25k;

static thread_local std::string g_err;
static std::map<std::string, double> g_kernel_ms;
static std::mutex g_ms_mu;

#define HIPCHK(x)                                                                   \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                     \
            g_err = std::string(#x) + ": " + hipGetErrorString(e_);                 \
            return F25_ERR_HIP;                                                     \
        }                                                                           \
    } while (0)
#define FAIL(code, msg)     \
    do {                    \
        g_err = (msg);      \
        return (code);      \
    } while (0)
#define TRY(x)                  \
    do {                        \
        int r_ = (x);           \
        if (r_) return r_;      \
    } while (0)

template <typename T>
static int dalloc(T **p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    HIPCHK(hipMalloc((void **)p, n * sizeof(T)));
    return 0;
}
template <typename T>
static int upload(T **p, const T *src, size_t n, hipStream_t s) {
    TRY(dalloc(p, n));
    if (n) HIPCHK(hipMemcpyAsync(*p, src, n * sizeof(T), hipMemcpyHostToDevice, s));
    return 0;
}
template <typename T>
static void dfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}
template <typename T>
static void fill_async(T *p, size_t n, T v, hipStream_t s) {
    if (!n) return;
    int nb = (int)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_fill<T>, dim3(nb), dim3(256), 0, s, p, n, v);
}

// a stream owned by one handle-free call
struct TmpStream {
    hipStream_t s = nullptr;
    ~TmpStream() { if (s) (void)hipStreamDestroy(s); }
};

struct f25_mesh_s {
    // host copies of the description
    f25_mesh_desc d{};
    std::vector<int> ieng;
    std::vector<double> xl, zl, xd, zd;
    std::string cdom, cnnpg;
    Tables tab;
    Graph g;
    InvInfo inv;
    bool have_inv = false;
    std::vector<int> order, cptr, pml_order, pml_cptr;
    std::vector<double> rho, lam, mu;
    bool have_model = false;
    // dispersive medium (quality factors at anchors, reference frequency)
    std::vector<double> qp, qs;
    double freq0 = 0.0;
    bool disp = false;
    double *d_qp = nullptr, *d_qs = nullptr;
    hipStream_t st = nullptr;
    int dev = 0;
    // device
    int *d_ieng = nullptr, *d_lm = nullptr, *d_order = nullptr, *d_pml_order = nullptr, *d_err = nullptr;
    unsigned long long *d_first = nullptr;   // first-touch node masks per element
    double *d_xl = nullptr, *d_zl = nullptr, *d_xd = nullptr, *d_zd = nullptr;
    double *d_shl = nullptr, *d_wq = nullptr, *d_xq = nullptr;
    double *d_t1d = nullptr;    // 1-D tables: val1 (nl*nq1) then der1
    char *d_cdom = nullptr, *d_cnp = nullptr;
    double *d_rho = nullptr, *d_lam = nullptr, *d_mu = nullptr;
    double *d_blk = nullptr;
    float *d_blk32 = nullptr;                          // fl32 of the blocks (FP32 mode)
    float *d_blk_hi = nullptr, *d_blk_lo = nullptr;   // fixed-point split (refined mode)
    unsigned long long *d_colmax = nullptr;           // per-pair max |x| (refined mode)
    int colmax_cap = 0;
    double *d_dblk = nullptr;   // blocks of a model perturbation (GN products)
    double *d_vcf = nullptr;    // d_blk in the coefficient layout of the FP64 product (k_apply_v)
    double *d_dvcf = nullptr;   // d_dblk in that layout
    size_t vcf_cap = 0, dvcf_cap = 0;
    // captured FP64 product (F25_APPLY_GRAPH) and the arguments it was captured for
    struct {
        hipGraphExec_t exec = nullptr;
        const void *x = nullptr, *vcf = nullptr, *om = nullptr, *kv = nullptr, *pad = nullptr;
        void *y = nullptr;
        int npair = 0;
        // the per-omega launches captured with it
        const void *mats = nullptr, *grp = nullptr, *gpair = nullptr;
        int nom = 0, ngrp = 0;
        bool disp = false;
    } ag;
    // PML elements in the batched product: per distinct omega of the batch
    // their element matrices by power of ky (k_pml_split), applied after the
    // element-block product by k_apply_pml over groups of pairs of one omega
    struct {
        std::vector<double> om;            // distinct omegas the matrices hold
        bool valid = false;                // cleared when the model or the medium changes
        bool disp = false;                 // the matrices cover every element (dispersive medium)
        double2 *d_mats = nullptr;         // [slot][omega][3][NEE][NEE]
        size_t mats_cap = 0;
        double *d_om = nullptr;
        int om_cap = 0;
        std::vector<int> grp, gpair;       // (omega index, first, count) per group; pairs by omega
        int *d_grp = nullptr, *d_gpair = nullptr;
        int grp_cap = 0, gpair_cap = 0, ngrp = 0;
    } pa;
    int apply_nrt = 0;          // row tiles of k_apply_v (0: the order's default)
    double *d_vpad = nullptr;   // k_apply_v: vpad_cap zero double2, then vpad_cap double2 of store sink
    int vpad_cap = 0;
    int64_t *d_irptr = nullptr;
    int *d_jcptr = nullptr;
    signed char *d_comp = nullptr;
    double *d_v0 = nullptr, *d_v1 = nullptr, *d_v2 = nullptr;   // V1, V2 compact (kind-0 entries)
    signed char *d_kind = nullptr;   // CSR value kind per entry (k_csr_kinds)
    int64_t *d_cbase = nullptr;      // kind-0 entries before each chunk of 64 entries
    int64_t n1 = 0;
    bool basis_ok = false;
    double *d_part = nullptr;
    int *d_mconn = nullptr;
    int *d_pmap = nullptr;   // pseudo-Hessian patch positions
    rocblas_handle blas = nullptr;   // the Gauss-Newton Hessian's SYRK
    // anchors of PML elements: parameters held fixed by the gradient-type
    // products (the reference inverts only anchors of interior elements)
    std::vector<char> fixa;
    char *d_fixa = nullptr;
    int ncon_anchor = 0;
    // inversion (CFMAT/ADJFLAM)
    int *d_mygrad = nullptr, *d_jcsc = nullptr, *d_icsc = nullptr;
    int *d_dab_pos = nullptr;   // [column][row of its last element]: position in ICSC_FDIST, -1 fixed
    double *d_ew = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // (omega, ky) staging, grown on demand
    double *d_omky = nullptr;
    int omky_cap = 0;
    // Bielak sources in the batch: the ring of elements holding both 'B' and
    // 'E' DOFs, by colour, with DOF maps selecting the input and output
    // classes of ASMB_PEFF's two terms, and the DOF positions of
    // FILL_UWAVE_QMESH; built on first use
    struct {
        bool built = false, uw = false;
        std::vector<int> cptr;
        int *d_order = nullptr;
        int *d_lm[4] = {};                       // E in, B out, B in, E out
        unsigned long long *d_first0 = nullptr;  // every row adds
        int *d_brows = nullptr;
        int nb = 0;
        double2 *d_xz = nullptr;
        signed char *d_comp = nullptr;
        double *d_pw = nullptr;
        int pw_cap = 0;
        // captured plane waves + effective forces and what they were captured for
        hipGraphExec_t exec = nullptr;
        const void *ue = nullptr, *peff = nullptr, *om = nullptr, *pw = nullptr, *pad = nullptr, *vcf = nullptr;
        int npair = 0;
    } br;
};

extern "C" const char *f25_last_error(void) { return g_err.c_str(); }

extern "C" double f25_last_kernel_ms(const char *family) {
    std::lock_guard<std::mutex> lk(g_ms_mu);
    auto it = g_kernel_ms.find(family ? family : "");
    return it == g_kernel_ms.end() ? -1.0 : it->second;
}

static void record_ms(const char *fam, double ms) {
    std::lock_guard<std::mutex> lk(g_ms_mu);
    g_kernel_ms[fam] = ms;
}

static int check_err(f25_mesh *m) {
    int h = 0;
    HIPCHK(hipMemcpyAsync(&h, m->d_err, sizeof(int), hipMemcpyDeviceToHost, m->st));
    HIPCHK(hipStreamSynchronize(m->st));
    if (h) {
        int z = 0;
        HIPCHK(hipMemcpy(m->d_err, &z, sizeof(int), hipMemcpyHostToDevice));
        if (h == 2) FAIL(F25_ERR_JACOBIAN, "Jacobian undefined or negative");
        if (h == 3) FAIL(F25_ERR_PATTERN, "element (row, col) pair missing from the CSR pattern");
        FAIL(F25_ERR_HIP, "device error flag");
    }
    return 0;
}

static int upload_mesh(f25_mesh *m) {
    const f25_mesh_desc &d = m->d;
    hipStream_t s = m->st;
    const int nen = m->tab.nen;
    TRY(upload(&m->d_ieng, m->ieng.data(), m->ieng.size(), s));
    TRY(upload(&m->d_xl, m->xl.data(), m->xl.size(), s));
    TRY(upload(&m->d_zl, m->zl.data(), m->zl.size(), s));
    TRY(upload(&m->d_xd, m->xd.data(), m->xd.size(), s));
    TRY(upload(&m->d_zd, m->zd.data(), m->zd.size(), s));
    TRY(upload(&m->d_cdom, m->cdom.data(), m->cdom.size(), s));
    TRY(upload(&m->d_shl, m->tab.shl.data(), m->tab.shl.size(), s));
    TRY(upload(&m->d_wq, m->tab.wq.data(), m->tab.wq.size(), s));
    TRY(upload(&m->d_xq, m->tab.xq.data(), m->tab.xq.size(), s));
    {
        std::vector<double> t1(m->tab.val1);
        t1.insert(t1.end(), m->tab.der1.begin(), m->tab.der1.end());
        TRY(upload(&m->d_t1d, t1.data(), t1.size(), s));
    }
    std::vector<int> lm0(m->g.lm.size());
    for (size_t i = 0; i < lm0.size(); ++i) lm0[i] = m->g.lm[i] - 1;   // 0-based, -1 fixed
    TRY(upload(&m->d_lm, lm0.data(), lm0.size(), s));
    TRY(upload(&m->d_order, m->order.data(), m->order.size(), s));
    {   // first touch: the element of lowest colour holding a node writes its y
        std::vector<int> colour(d.nelem), mincol(m->g.ndof + 1, 1 << 30);
        for (size_t c = 0; c + 1 < m->cptr.size(); ++c)
            for (int i = m->cptr[c]; i < m->cptr[c + 1]; ++i) colour[m->order[i]] = (int)c;
        auto key = [&](int e, int a) {
            const int *l = &m->g.lm[((size_t)e * nen + a) * 3];
            return l[0] > 0 ? l[0] : (l[1] > 0 ? l[1] : l[2]);
        };
        for (int e = 0; e < d.nelem; ++e)
            for (int a = 0; a < nen; ++a) {
                int k = key(e, a);
                if (k > 0) mincol[k] = std::min(mincol[k], colour[e]);
            }
        std::vector<unsigned long long> first(d.nelem, 0ull);
        for (int e = 0; e < d.nelem; ++e)
            for (int a = 0; a < nen; ++a) {
                int k = key(e, a);
                if (k > 0 && mincol[k] == colour[e]) first[e] |= 1ull << a;
            }
        TRY(upload(&m->d_first, first.data(), first.size(), s));
    }
    TRY(upload(&m->d_pml_order, m->pml_order.data(), m->pml_order.size(), s));
    TRY(dalloc(&m->d_err, 1));
    HIPCHK(hipMemsetAsync(m->d_err, 0, sizeof(int), s));
    // anchor connectivity for the gradient gather
    std::vector<int> cnt(d.nnpg, 0);
    for (int e = 0; e < d.nelem; ++e)
        for (int k = 0; k < 4; ++k) cnt[m->ieng[4 * e + k] - 1]++;
    m->ncon_anchor = *std::max_element(cnt.begin(), cnt.end());
    std::vector<int> mc((size_t)m->ncon_anchor * d.nnpg, 0), f(d.nnpg, 0);
    for (int e = 0; e < d.nelem; ++e)
        for (int k = 0; k < 4; ++k) {
            int a = m->ieng[4 * e + k] - 1;
            mc[(size_t)f[a]++ * d.nnpg + a] = e + 1;
        }
    TRY(upload(&m->d_mconn, mc.data(), mc.size(), s));
    m->fixa.assign(d.nnpg, 0);
    for (int e = 0; e < d.nelem; ++e)
        if (m->cdom[e] == 'A')
            for (int k = 0; k < 4; ++k) m->fixa[m->ieng[4 * e + k] - 1] = 1;
    TRY(upload(&m->d_fixa, m->fixa.data(), m->fixa.size(), s));
    TRY(dalloc(&m->d_part, (size_t)d.nelem * 12));
    (void)nen;
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

static int init_handle(f25_mesh *m, const f25_mesh_desc *desc) {
    if (!desc || desc->nord < 2 || desc->nord > 6 || desc->nelem <= 0 || desc->nnpg <= 0 ||
        !desc->ieng || !desc->xlocs || !desc->zlocs || !desc->cdomain || !desc->cnnpg)
        FAIL(F25_ERR_ARG, "invalid mesh description (NORD must be 2..6)");
    m->d = *desc;
    m->ieng.assign(desc->ieng, desc->ieng + 4 * (size_t)desc->nelem);
    for (int v : m->ieng)
        if (v < 1 || v > desc->nnpg) FAIL(F25_ERR_ARG, "IENG entry out of range");
    m->xl.assign(desc->xlocs, desc->xlocs + desc->nnpg);
    m->zl.assign(desc->zlocs, desc->zlocs + desc->nnpg);
    m->xd.assign(desc->nnpg, 0.0);
    m->zd.assign(desc->nnpg, 0.0);
    if (desc->xd) m->xd.assign(desc->xd, desc->xd + desc->nnpg);
    if (desc->zd) m->zd.assign(desc->zd, desc->zd + desc->nnpg);
    m->cdom.assign(desc->cdomain, desc->nelem);
    m->cnnpg.assign(desc->cnnpg, 2 * (size_t)desc->nnpg);
    m->d.ieng = m->ieng.data();
    m->d.xlocs = m->xl.data();
    m->d.zlocs = m->zl.data();
    m->d.xd = m->xd.data();
    m->d.zd = m->zd.data();
    m->d.cdomain = m->cdom.data();
    m->d.cnnpg = m->cnnpg.data();
    make_tables(desc->nord, m->tab);
    HIPCHK(hipGetDevice(&m->dev));
    // a blocking stream: it orders with work on the legacy default stream, which
    // is where callers such as PyTorch produce the device inputs
    HIPCHK(hipStreamCreate(&m->st));
    HIPCHK(hipEventCreate(&m->ev0));
    HIPCHK(hipEventCreate(&m->ev1));
    // colouring: all elements, and the PML elements on their own
    colour_elements(desc->nelem, desc->nnpg, m->ieng.data(), m->order, m->cptr);
    std::vector<int> pmle, pieng;
    for (int e = 0; e < desc->nelem; ++e)
        if (m->cdom[e] == 'A') pmle.push_back(e);
    if (!pmle.empty()) {
        for (int e : pmle)
            for (int k = 0; k < 4; ++k) pieng.push_back(m->ieng[4 * e + k]);
        std::vector<int> ord;
        colour_elements((int)pmle.size(), desc->nnpg, pieng.data(), ord, m->pml_cptr);
        m->pml_order.resize(ord.size());
        for (size_t i = 0; i < ord.size(); ++i) m->pml_order[i] = pmle[ord[i]];
    } else {
        m->pml_cptr.assign(1, 0);
    }
    return 0;
}

// ---------------------------------------------------------------------------
// host-only numbering
struct f25_graph_s {
    Tables tab;
    Graph g;
    int nelem = 0;
};

extern "C" int f25_graph_build(const f25_mesh_desc *desc, int with_csr, f25_graph **out,
                               int64_t sizes[3]) {
    if (!out || !desc || desc->nord < 2 || desc->nord > 6 || desc->nelem <= 0 || desc->nnpg <= 0 ||
        !desc->ieng || !desc->xlocs || !desc->zlocs || !desc->cdomain || !desc->cnnpg)
        FAIL(F25_ERR_ARG, "invalid mesh description (NORD must be 2..6)");
    *out = nullptr;
    for (size_t i = 0; i < 4 * (size_t)desc->nelem; ++i)
        if (desc->ieng[i] < 1 || desc->ieng[i] > desc->nnpg) FAIL(F25_ERR_ARG, "IENG entry out of range");
    std::unique_ptr<f25_graph> g(new f25_graph_s());
    g->nelem = desc->nelem;
    make_tables(desc->nord, g->tab);
    int r = build_numbering(*desc, g->tab, g->g);
    if (r) FAIL(r, "GLL numbering failed");
    if (with_csr) {
        r = build_csr(*desc, g->tab, g->g);
        if (r) FAIL(r, "CSR pattern build failed");
    }
    if (sizes) {
        sizes[0] = g->g.nnp;
        sizes[1] = g->g.ndof;
        sizes[2] = g->g.have_csr ? g->g.irptr[g->g.ndof] : 0;
    }
    *out = g.release();
    return 0;
}

extern "C" int f25_graph_get(const f25_graph *g, int *ien, int *lm, int *irptr, int *jcptr, char *cnp) {
    if (!g) FAIL(F25_ERR_ARG, "null graph");
    if ((irptr || jcptr) && !g->g.have_csr) FAIL(F25_ERR_STATE, "graph built without CSR pattern");
    if (ien) std::copy(g->g.ien.begin(), g->g.ien.end(), ien);
    if (lm) std::copy(g->g.lm.begin(), g->g.lm.end(), lm);
    if (irptr)
        for (int i = 0; i <= g->g.ndof; ++i) irptr[i] = (int)(g->g.irptr[i] + 1);
    if (jcptr) std::copy(g->g.jcptr.begin(), g->g.jcptr.end(), jcptr);
    if (cnp) std::copy(g->g.cnp.begin(), g->g.cnp.end(), cnp);
    return 0;
}

extern "C" int f25_graph_destroy(f25_graph *g) {
    delete g;
    return 0;
}

extern "C" int f25_mesh_create(const f25_mesh_desc *desc, f25_mesh **out) {
    if (!out) FAIL(F25_ERR_ARG, "null output handle");
    *out = nullptr;
    std::unique_ptr<f25_mesh> m(new f25_mesh_s());
    TRY(init_handle(m.get(), desc));
    int r = build_numbering(m->d, m->tab, m->g);
    if (r) FAIL(r, "GLL numbering failed");
    TRY(upload_mesh(m.get()));
    *out = m.release();
    return 0;
}

// A caller's 1-based CSR pattern: IRPTR(1) = 1, ascending row pointers,
// column numbers in 1..ncol (the device kernels index x and the output by
// them without further checks).
static int check_csr(int nrow, int ncol, const int *irptr, const int *jcptr) {
    if (irptr[0] != 1) FAIL(F25_ERR_ARG, "IRPTR(1) must be 1");
    for (int i = 0; i < nrow; ++i)
        if (irptr[i + 1] < irptr[i]) FAIL(F25_ERR_ARG, "IRPTR not ascending");
    const int64_t nz = (int64_t)irptr[nrow] - 1;
    for (int64_t k = 0; k < nz; ++k)
        if (jcptr[k] < 1 || jcptr[k] > ncol) FAIL(F25_ERR_ARG, "JCPTR entry out of range");
    return 0;
}

extern "C" int f25_mesh_create_with_graph(const f25_mesh_desc *desc, int ndof, const int *lm,
                                          const int *irptr, const int *jcptr, f25_mesh **out) {
    if (!out || !lm || ndof <= 0) FAIL(F25_ERR_ARG, "invalid graph arguments");
    *out = nullptr;
    std::unique_ptr<f25_mesh> m(new f25_mesh_s());
    TRY(init_handle(m.get(), desc));
    const int nen = m->tab.nen;
    m->g.ndof = ndof;
    m->g.lm.assign(lm, lm + (size_t)desc->nelem * nen * 3);
    for (int v : m->g.lm)
        if (v < 0 || v > ndof) FAIL(F25_ERR_ARG, "LM entry out of range");
    m->g.nnp = 0;
    m->g.cnp.assign(ndof, ' ');
    for (char dom : {'A', 'E', 'I'})
        for (int e = 0; e < desc->nelem; ++e) {
            if (m->cdom[e] != dom) continue;
            for (int k = 0; k < 3 * nen; ++k) {
                int dd = m->g.lm[(size_t)e * 3 * nen + k];
                if (dd <= 0) continue;
                char &c = m->g.cnp[dd - 1];
                if (dom == 'A') c = 'A';
                else if (dom == 'E') { if (c == ' ') c = 'E'; }
                else if (c == ' ') c = 'I';
                else if (c == 'E') c = 'B';
            }
        }
    if (irptr) {
        if (!jcptr) FAIL(F25_ERR_ARG, "JCPTR missing with IRPTR");
        TRY(check_csr(ndof, ndof, irptr, jcptr));
        m->g.irptr.resize(ndof + 1);
        for (int i = 0; i <= ndof; ++i) m->g.irptr[i] = (int64_t)irptr[i] - 1;
        m->g.jcptr.assign(jcptr, jcptr + m->g.irptr[ndof]);
        m->g.have_csr = true;
    }
    TRY(upload_mesh(m.get()));
    *out = m.release();
    return 0;
}

extern "C" int f25_mesh_destroy(f25_mesh *m) {
    if (!m) return 0;
    (void)hipSetDevice(m->dev);
    if (m->st) (void)hipStreamSynchronize(m->st);
    dfree(m->d_ieng); dfree(m->d_lm); dfree(m->d_first); dfree(m->d_order); dfree(m->d_pml_order); dfree(m->d_err);
    dfree(m->d_xl); dfree(m->d_zl); dfree(m->d_xd); dfree(m->d_zd); dfree(m->d_shl); dfree(m->d_wq); dfree(m->d_t1d);
    dfree(m->d_xq); dfree(m->d_cdom); dfree(m->d_cnp); dfree(m->d_rho); dfree(m->d_lam); dfree(m->d_mu);
    if (m->ag.exec) (void)hipGraphExecDestroy(m->ag.exec);
    dfree(m->d_vcf); dfree(m->d_dvcf); dfree(m->d_vpad);
    dfree(m->d_blk); dfree(m->d_blk32); dfree(m->d_blk_hi); dfree(m->d_blk_lo); dfree(m->d_colmax); dfree(m->d_dblk); dfree(m->d_irptr);
    dfree(m->d_jcptr); dfree(m->d_comp); dfree(m->d_v0); dfree(m->d_v1); dfree(m->d_v2);
    dfree(m->d_kind); dfree(m->d_cbase);
    dfree(m->d_part); dfree(m->d_mconn); dfree(m->d_pmap); dfree(m->d_fixa); dfree(m->d_qp); dfree(m->d_qs);
    dfree(m->d_mygrad); dfree(m->d_jcsc); dfree(m->d_icsc); dfree(m->d_dab_pos);
    dfree(m->d_ew);
    dfree(m->d_omky);
    dfree(m->br.d_order); dfree(m->br.d_first0); dfree(m->br.d_brows); dfree(m->br.d_xz); dfree(m->br.d_comp);
    dfree(m->br.d_pw);
    for (auto &q : m->br.d_lm) dfree(q);
    if (m->br.exec) (void)hipGraphExecDestroy(m->br.exec);
    dfree(m->pa.d_mats); dfree(m->pa.d_om); dfree(m->pa.d_grp); dfree(m->pa.d_gpair);
    if (m->blas) (void)rocblas_destroy_handle(m->blas);
    if (m->ev0) (void)hipEventDestroy(m->ev0);
    if (m->ev1) (void)hipEventDestroy(m->ev1);
    if (m->st) (void)hipStreamDestroy(m->st);
    delete m;
    return 0;
}

static int ensure_csr(f25_mesh *m) {
    if (!m->g.have_csr) {
        int r = build_csr(m->d, m->tab, m->g);
        if (r) FAIL(r, "CSR pattern build failed");
    }
    if (!m->d_irptr) {
        hipStream_t s = m->st;
        std::vector<int> jc0(m->g.jcptr.size());
        for (size_t i = 0; i < jc0.size(); ++i) jc0[i] = m->g.jcptr[i] - 1;
        TRY(upload(&m->d_irptr, m->g.irptr.data(), m->g.irptr.size(), s));
        TRY(upload(&m->d_jcptr, jc0.data(), jc0.size(), s));
        std::vector<signed char> comp(m->g.ndof, 0);
        for (size_t i = 0; i < m->g.lm.size(); ++i)
            if (m->g.lm[i] > 0) comp[m->g.lm[i] - 1] = (signed char)(i % 3);
        TRY(upload(&m->d_comp, comp.data(), comp.size(), s));
        TRY(upload(&m->d_cnp, m->g.cnp.data(), m->g.cnp.size(), s));
        HIPCHK(hipStreamSynchronize(s));
    }
    return 0;
}

extern "C" int f25_mesh_sizes(f25_mesh *m, int64_t sizes[6]) {
    if (!m || !sizes) FAIL(F25_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(m->dev));
    TRY(ensure_csr(m));
    sizes[0] = m->g.nnp;
    sizes[1] = m->g.ndof;
    sizes[2] = m->g.irptr.empty() ? 0 : m->g.irptr[m->g.ndof];
    sizes[3] = (int64_t)m->cptr.size() - 1;
    sizes[4] = m->tab.nen;
    sizes[5] = m->tab.nq1;
    return 0;
}

extern "C" int f25_mesh_graph(f25_mesh *m, int *ien, int *lm, int *irptr, int *jcptr, char *cnp) {
    if (!m) FAIL(F25_ERR_ARG, "null handle");
    if ((irptr || jcptr)) {
        HIPCHK(hipSetDevice(m->dev));
        TRY(ensure_csr(m));
    }
    if (ien && !m->g.ien.empty()) std::copy(m->g.ien.begin(), m->g.ien.end(), ien);
    if (lm) std::copy(m->g.lm.begin(), m->g.lm.end(), lm);
    if (irptr)
        for (int i = 0; i <= m->g.ndof; ++i) irptr[i] = (int)(m->g.irptr[i] + 1);
    if (jcptr) std::copy(m->g.jcptr.begin(), m->g.jcptr.end(), jcptr);
    if (cnp) std::copy(m->g.cnp.begin(), m->g.cnp.end(), cnp);
    return 0;
}

// ---------------------------------------------------------------------------
template <int NL>
static void launch_blocks(f25_mesh *m, const double *rho, const double *lam, const double *mu,
                          double *blk, float *b32, float *hi, float *lo) {
    constexpr int NT = (NL <= 5) ? 256 : 512;
    hipLaunchKernelGGL((k_blocks<NL, NT>), dim3(m->d.nelem), dim3(NT), 0, m->st, m->d.nelem,
                       m->tab.nq1, m->d_shl, m->d_wq, m->d_xq, m->d_ieng, m->d_xl, m->d_zl, rho, lam,
                       mu, m->d_cdom, blk, b32, hi, lo, m->d_err);
}

static int run_blocks(f25_mesh *m, const double *rho, const double *lam, const double *mu,
                      double *blk, float *b32, float *hi, float *lo) {
    switch (m->tab.nl) {
        case 3: launch_blocks<3>(m, rho, lam, mu, blk, b32, hi, lo); break;
        case 4: launch_blocks<4>(m, rho, lam, mu, blk, b32, hi, lo); break;
        case 5: launch_blocks<5>(m, rho, lam, mu, blk, b32, hi, lo); break;
        case 6: launch_blocks<6>(m, rho, lam, mu, blk, b32, hi, lo); break;
        case 7: launch_blocks<7>(m, rho, lam, mu, blk, b32, hi, lo); break;
        default: FAIL(F25_ERR_ARG, "unsupported order");
    }
    HIPCHK(hipGetLastError());
    return check_err(m);
}

static int vcoef(f25_mesh *m, const double *blk, double **vcf);

extern "C" int f25_set_model(f25_mesh *m, const double *dens, const double *ecoeff) {
    if (!m || !dens || !ecoeff) FAIL(F25_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(m->dev));
    const int n = m->d.nnpg;
    // unchanged model: the element blocks on the device are current
    if (m->have_model && std::equal(dens, dens + n, m->rho.begin()) &&
        std::equal(ecoeff, ecoeff + n, m->lam.begin()) && std::equal(ecoeff + n, ecoeff + 2 * n, m->mu.begin()))
        return 0;
    m->have_model = false;
    m->rho.assign(dens, dens + n);
    m->lam.assign(ecoeff, ecoeff + n);
    m->mu.assign(ecoeff + n, ecoeff + 2 * n);
    for (int i = 0; i < n; ++i)
        if (!(m->rho[i] > 0.0)) FAIL(F25_ERR_ARG, "density must be positive");
    if (!m->d_rho) {
        TRY(dalloc(&m->d_rho, n));
        TRY(dalloc(&m->d_lam, n));
        TRY(dalloc(&m->d_mu, n));
    }
    HIPCHK(hipMemcpyAsync(m->d_rho, m->rho.data(), n * 8, hipMemcpyHostToDevice, m->st));
    HIPCHK(hipMemcpyAsync(m->d_lam, m->lam.data(), n * 8, hipMemcpyHostToDevice, m->st));
    HIPCHK(hipMemcpyAsync(m->d_mu, m->mu.data(), n * 8, hipMemcpyHostToDevice, m->st));
    const size_t nb = (size_t)m->d.nelem * NBLK * m->tab.nen * m->tab.nen;
    if (!m->d_blk) {
        TRY(dalloc(&m->d_blk, nb));
        TRY(dalloc(&m->d_blk32, nb));
        TRY(dalloc(&m->d_blk_hi, nb));
        TRY(dalloc(&m->d_blk_lo, nb));
    }
    TRY(run_blocks(m, m->d_rho, m->d_lam, m->d_mu, m->d_blk, m->d_blk32, m->d_blk_hi, m->d_blk_lo));
    TRY(vcoef(m, m->d_blk, &m->d_vcf));
    m->have_model = true;
    m->basis_ok = false;
    m->pa.valid = false;
    return 0;
}

extern "C" int f25_set_attenuation(f25_mesh *m, const double *qp, const double *qs, double freq0) {
    if (!m) FAIL(F25_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(m->dev));
    if (!qp || !qs) {
        if (m->disp) m->pa.valid = false;
        m->disp = false;
        return 0;
    }
    const int n = m->d.nnpg;
    if (m->disp && freq0 == m->freq0 && std::equal(qp, qp + n, m->qp.begin()) &&
        std::equal(qs, qs + n, m->qs.begin()))
        return 0;
    m->qp.assign(qp, qp + n);
    m->qs.assign(qs, qs + n);
    m->freq0 = freq0;
    if (!m->d_qp) {
        TRY(dalloc(&m->d_qp, n));
        TRY(dalloc(&m->d_qs, n));
    }
    HIPCHK(hipMemcpyAsync(m->d_qp, m->qp.data(), n * 8, hipMemcpyHostToDevice, m->st));
    HIPCHK(hipMemcpyAsync(m->d_qs, m->qs.data(), n * 8, hipMemcpyHostToDevice, m->st));
    HIPCHK(hipStreamSynchronize(m->st));
    m->disp = true;
    m->pa.valid = false;
    return 0;
}

// entry points built on the frequency-independent element blocks
static int require_elastic(f25_mesh *m) {
    if (m->disp) FAIL(F25_ERR_STATE, "dispersive media are supported by the assembled operator only");
    return 0;
}

// ---------------------------------------------------------------------------
// value kinds of the CSR entries and the chunk bases of the compact V1, V2
static int ensure_kinds(f25_mesh *m) {
    if (m->d_kind) return 0;
    const int ndof = m->g.ndof;
    const int64_t nz = m->g.irptr[ndof];
    const int64_t nchunk = (nz + 63) / 64;
    if (nchunk / 4 + 1 > 0x7fffffff) FAIL(F25_ERR_ARG, "too many CSR entries");
    int *d_cnt = nullptr;
    TRY(dalloc(&m->d_kind, (size_t)std::max<int64_t>(nz, 1)));
    TRY(dalloc(&d_cnt, (size_t)std::max<int64_t>(nchunk, 1)));
    hipLaunchKernelGGL(k_csr_kinds, dim3((ndof + 3) / 4), dim3(256), 0, m->st, ndof, m->d_irptr, m->d_jcptr,
                       m->d_comp, m->d_kind);
    if (nchunk > 0)
        hipLaunchKernelGGL(k_csr_chunks, dim3((unsigned)((nchunk + 3) / 4)), dim3(256), 0, m->st, nz, m->d_kind,
                           d_cnt);
    std::vector<int> cnt(nchunk);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && nchunk > 0)
        e = hipMemcpyAsync(cnt.data(), d_cnt, (size_t)nchunk * sizeof(int), hipMemcpyDeviceToHost, m->st);
    if (e == hipSuccess) e = hipStreamSynchronize(m->st);
    dfree(d_cnt);
    if (e != hipSuccess) FAIL(F25_ERR_HIP, hipGetErrorString(e));
    std::vector<int64_t> base(nchunk + 1, 0);
    for (int64_t q = 0; q < nchunk; ++q) base[q + 1] = base[q] + cnt[q];
    m->n1 = base[nchunk];
    TRY(upload(&m->d_cbase, base.data(), base.size(), m->st));
    HIPCHK(hipStreamSynchronize(m->st));
    return 0;
}

// grid of a chunk-stream kernel: enough workgroups of 4 waves to fill the
// chip several times over, each wave then striding over the chunks
static unsigned chunk_grid(int64_t nz) {
    const int64_t nchunk = (nz + 127) / 128;   // k_csr_combine: 128 entries per wave step
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((nchunk + 3) / 4, 8192));
}

static int ensure_basis(f25_mesh *m) {
    TRY(ensure_csr(m));
    TRY(ensure_kinds(m));
    if (m->basis_ok) return 0;
    const int ndof = m->g.ndof;
    const size_t nz = m->g.irptr[ndof];
    if (!m->d_v0) {
        TRY(dalloc(&m->d_v0, nz));
        TRY(dalloc(&m->d_v1, (size_t)std::max<int64_t>(m->n1, 1)));
        TRY(dalloc(&m->d_v2, (size_t)std::max<int64_t>(m->n1, 1)));
    }
    fill_async(m->d_v0, nz, 0.0, m->st);
    fill_async(m->d_v1, (size_t)m->n1, 0.0, m->st);
    fill_async(m->d_v2, (size_t)m->n1, 0.0, m->st);
    // compact positions, needed only while the basis is scattered
    int64_t *d_cidx = nullptr;
    TRY(dalloc(&d_cidx, nz));
    if (nz > 0)
        hipLaunchKernelGGL(k_csr_cidx, dim3((unsigned)((((int64_t)nz + 63) / 64 + 3) / 4)), dim3(256), 0, m->st,
                           (int64_t)nz, m->d_cbase, m->d_kind, d_cidx);
    for (size_t c = 0; c + 1 < m->cptr.size(); ++c) {
        int n = m->cptr[c + 1] - m->cptr[c];
        if (!n) continue;
        hipLaunchKernelGGL(k_csr_basis, dim3(n), dim3(256), 0, m->st, m->tab.nen, n,
                           m->d_order + m->cptr[c], m->d_blk, m->d_lm, m->d_irptr, m->d_jcptr, d_cidx,
                           m->d_cdom, m->d_v0, m->d_v1, m->d_v2, m->d_err);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(m->st);
    dfree(d_cidx);
    if (e != hipSuccess) FAIL(F25_ERR_HIP, hipGetErrorString(e));
    TRY(check_err(m));
    m->basis_ok = true;
    return 0;
}

template <typename CT>
static int asmble_impl(f25_mesh *m, char cftype, double freq, double py, CT *out) {
    const double omega = 6.2831853071795862 * freq, kv = omega * py;
    const int ndof = m->g.ndof;
    HIPCHK(hipEventRecord(m->ev0, m->st));
    // dispersive media: every element is integrated per call (KMDAMP/CMDAMP);
    // otherwise the frequency-independent basis gives the non-PML part
    const bool disp = m->disp;
    if (disp)
        fill_async(out, (size_t)m->g.irptr[ndof], CT{0, 0}, m->st);
    else
        hipLaunchKernelGGL((k_csr_combine<CT>), dim3(chunk_grid(m->g.irptr[ndof])), dim3(256), 0, m->st,
                           (int64_t)m->g.irptr[ndof], m->d_cbase, m->d_kind, m->d_v0, m->d_v1, m->d_v2, omega, kv,
                           out);
    const std::vector<int> &cp = disp ? m->cptr : m->pml_cptr;
    const int *ord = disp ? m->d_order : m->d_pml_order;
    if (disp || !m->pml_order.empty()) {
        PmlPar p;
        p.omega = omega;
        p.kv = kv;
        p.xwidth = m->d.xwidth;
        p.zwidth = m->d.zwidth;
        const double *pp = (cftype == 'S') ? m->d.pml_srf : m->d.pml;
        p.rcoeff = pp[0]; p.dmax = pp[1]; p.kappa = pp[2]; p.fcent = pp[3];
        p.freq = freq;
        p.freq0 = m->freq0;
        const int nen = m->tab.nen;
        size_t lds = (8 * 9 + 3 * 8 * nen) * sizeof(double);
        for (size_t c = 0; c + 1 < cp.size(); ++c) {
            int n = cp[c + 1] - cp[c];
            if (!n) continue;
            hipLaunchKernelGGL((k_pml<CT>), dim3(n), dim3(256), lds, m->st, nen, m->tab.nq1, n,
                               ord + cp[c], m->d_shl, m->d_wq, m->d_xq, m->d_ieng,
                               m->d_xl, m->d_zl, m->d_xd, m->d_zd, m->d_cdom, m->d_rho, m->d_lam, m->d_mu,
                               disp ? m->d_qp : nullptr, disp ? m->d_qs : nullptr, m->d_lm,
                               m->d_irptr, m->d_jcptr, p, out, m->d_err);
        }
    }
    HIPCHK(hipEventRecord(m->ev1, m->st));
    HIPCHK(hipGetLastError());
    TRY(check_err(m));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, m->ev0, m->ev1));
    record_ms("asmble", ms);
    return 0;
}

extern "C" int f25_asmble_dev(f25_mesh *m, char cftype, double freq, double py, int precision, void *out) {
    if (!m || !out) FAIL(F25_ERR_ARG, "null argument");
    if (!m->have_model) FAIL(F25_ERR_STATE, "f25_set_model has not been called");
    HIPCHK(hipSetDevice(m->dev));
    TRY(ensure_basis(m));
    if (precision == F25_C128) return asmble_impl(m, cftype, freq, py, (double2 *)out);
    if (precision == F25_C64) return asmble_impl(m, cftype, freq, py, (float2 *)out);
    FAIL(F25_ERR_ARG, "precision must be F25_C64 or F25_C128");
}

extern "C" int f25_asmble(f25_mesh *m, char cftype, double freq, double py, int precision, void *out) {
    if (!m || !out) FAIL(F25_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(m->dev));
    TRY(ensure_csr(m));
    const size_t nz = m->g.irptr[m->g.ndof];
    const size_t bytes = nz * (precision == F25_C128 ? 16 : 8);
    void *d = nullptr;
    HIPCHK(hipMalloc(&d, bytes));
    int r = f25_asmble_dev(m, cftype, freq, py, precision, d);
    if (!r) {
        hipError_t e = hipMemcpyAsync(out, d, bytes, hipMemcpyDeviceToHost, m->st);
        if (e == hipSuccess) e = hipStreamSynchronize(m->st);
        if (e != hipSuccess) { g_err = hipGetErrorString(e); r = F25_ERR_HIP; }
    }
    (void)hipFree(d);
    return r;
}

extern "C" int f25_csr_apply_dev(f25_mesh *m, int precision, const void *s, const void *x, void *y) {
    if (!m || !s || !x || !y) FAIL(F25_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(m->dev));
    TRY(ensure_csr(m));
    const int ndof = m->g.ndof;
    HIPCHK(hipEventRecord(m->ev0, m->st));
    if (precision == F25_C128)
        hipLaunchKernelGGL((k_spmv<double2, double>), dim3((ndof + 3) / 4), dim3(256), 0, m->st, ndof,
                           m->d_irptr, m->d_jcptr, (const double2 *)s, (const double2 *)x, (double2 *)y);
    else if (precision == F25_C64)
        hipLaunchKernelGGL((k_spmv<float2, float>), dim3((ndof + 3) / 4), dim3(256), 0, m->st, ndof,
                           m->d_irptr, m->d_jcptr, (const float2 *)s, (const float2 *)x, (float2 *)y);
    else
        FAIL(F25_ERR_ARG, "bad precision");
    HIPCHK(hipEventRecord(m->ev1, m->st));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(m->st));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, m->ev0, m->ev1));
    record_ms("spmv", ms);
    return 0;
}

extern "C" int f25_asmb_peff(f25_mesh *m, const void *cdist, const void *ue, void *peff) {
    if (!m || !cdist || !ue || !peff) FAIL(F25_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(m->dev));
    TRY(ensure_csr(m));
    const int ndof = m->g.ndof;
    const size_t nz = m->g.irptr[ndof];
    float2 *ds = nullptr, *du = nullptr, *dp = nullptr;
    TRY(upload(&ds, (const float2 *)cdist, nz, m->st));
    TRY(upload(&du, (const float2 *)ue, ndof, m->st));
    TRY(dalloc(&dp, ndof));
    fill_async(dp, (size_t)ndof, make_float2(0.f, 0.f), m->st);
    hipLaunchKernelGGL(k_peff, dim3((ndof + 3) / 4), dim3(256), 0, m->st, ndof, (const int *)nullptr,
                       m->d_irptr, m->d_jcptr, m->d_cnp, ds, du, dp);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(peff, dp, ndof * sizeof(float2), hipMemcpyDeviceToHost, m->st);
    if (e == hipSuccess) e = hipStreamSynchronize(m->st);
    dfree(ds); dfree(du); dfree(dp);
    if (e != hipSuccess) FAIL(F25_ERR_HIP, hipGetErrorString(e));
    return 0;
}

// ---------------------------------------------------------------------------
template <int NL, typename BT, typename XT, bool REFINE, int CSEL>
static int launch_apply_c(f25_mesh *m, const double *om, const double *kv, int npair, const XT *x,
                          XT *y, const BT *blk, const BT *blk_lo) {
    const double *cmx = REFINE ? reinterpret_cast<const double *>(m->d_colmax) : nullptr;
    typedef ApplyLds<NL, BT, REFINE> LD;
    constexpr size_t lds = CSEL < 0 ? LD::bytes : (5 * LD::per_blk + 3 * ApplyShape<NL>::NEN * sizeof(int));
    static bool attr_set = false;
    auto kern = k_apply<NL, BT, XT, REFINE, CSEL>;
    if (!attr_set) {
        HIPCHK(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        attr_set = true;
    }
    for (size_t c = 0; c + 1 < m->cptr.size(); ++c) {
        int n = m->cptr[c + 1] - m->cptr[c];
        if (!n) continue;
        hipLaunchKernelGGL(kern, dim3(n), dim3(64 * ApplyLds<NL, BT, REFINE>::NW), lds, m->st,
                           m->d_order + m->cptr[c], blk, blk_lo,
                           m->d_lm, m->d_first, npair, om, kv, cmx, x, y);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

template <int NL, typename BT, typename XT, bool REFINE>
static int launch_apply(f25_mesh *m, const double *om, const double *kv, int npair, const XT *x,
                        XT *y, const BT *blk, const BT *blk_lo) {
    if constexpr (!ApplyLds<NL, BT, REFINE>::PH) {
        return launch_apply_c<NL, BT, XT, REFINE, -1>(m, om, kv, npair, x, y, blk, blk_lo);
    } else {
        // one output component per pass; the colour order is kept within each pass
        TRY((launch_apply_c<NL, BT, XT, REFINE, 0>(m, om, kv, npair, x, y, blk, blk_lo)));
        TRY((launch_apply_c<NL, BT, XT, REFINE, 1>(m, om, kv, npair, x, y, blk, blk_lo)));
        return launch_apply_c<NL, BT, XT, REFINE, 2>(m, om, kv, npair, x, y, blk, blk_lo);
    }
}

template <typename BT, typename XT, bool REFINE>
static int dispatch_apply(f25_mesh *m, const double *om, const double *kv, int npair, const XT *x,
                          XT *y, const BT *blk, const BT *blk_lo) {
    switch (m->tab.nl) {
        case 3: return launch_apply<3, BT, XT, REFINE>(m, om, kv, npair, x, y, blk, blk_lo);
        case 4: return launch_apply<4, BT, XT, REFINE>(m, om, kv, npair, x, y, blk, blk_lo);
        case 5: return launch_apply<5, BT, XT, REFINE>(m, om, kv, npair, x, y, blk, blk_lo);
        case 6: return launch_apply<6, BT, XT, REFINE>(m, om, kv, npair, x, y, blk, blk_lo);
        case 7: return launch_apply<7, BT, XT, REFINE>(m, om, kv, npair, x, y, blk, blk_lo);
    }
    FAIL(F25_ERR_ARG, "unsupported order");
}

// workgroups of one k_apply_v launch over n elements: per XCD (eight), its
// ceil(n / 8) elements times the pair groups times WPT waves, four waves
// per workgroup
template <int NL, int NRT>
static size_t apply_v_grid(int n, int npair) {
    const size_t ngrp = (size_t)(npair + 63) / 64, wpt = VShape<NL, NRT>::WPT;
    return 8 * (((size_t)(n + 7) / 8 * ngrp * wpt + 3) / 4);
}

// FP64 product (k_apply_v): one launch per element colour
template <int NL, int NRT>
static int launch_apply_v(f25_mesh *m, const double *om, const double *kv, int npair, const double2 *x,
                          double2 *y, const double *vcf) {
    const double2 *zpad = reinterpret_cast<const double2 *>(m->d_vpad);
    double2 *sink = reinterpret_cast<double2 *>(m->d_vpad) + m->vpad_cap;
    for (size_t c = 0; c + 1 < m->cptr.size(); ++c) {
        int n = m->cptr[c + 1] - m->cptr[c];
        if (!n) continue;
        const size_t nwg = apply_v_grid<NL, NRT>(n, npair);
        if (nwg > 0x7fffffff) FAIL(F25_ERR_ARG, "too many pairs for one launch");
        hipLaunchKernelGGL((k_apply_v<NL, NRT>), dim3((unsigned)nwg), dim3(VShape<NL, NRT>::NT), 0, m->st,
                           m->d_order + m->cptr[c], n, vcf, m->d_lm, m->d_lm, m->d_first, npair, om, kv, x, y,
                           zpad, sink);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

// Row tiles of the FP64 product per order (index NL = p + 1): the candidates
// compiled in, and the default, picked by fem25_amd/tune_apply.py on MI355X
// (400x200 elements, 192 pairs).
static bool apply_nrt_ok(int nl, int nrt) {
    switch (nl) {
        case 3: case 4: return nrt == 1 || nrt == 2;
        case 5: return nrt >= 1 && nrt <= 3;
        case 6: case 7: return nrt >= 2 && nrt <= 4;
    }
    return false;
}
static int apply_nrt_default(int nl) { return nl <= 5 ? 1 : 2; }
static int apply_nrt(const f25_mesh *m) { return m->apply_nrt ? m->apply_nrt : apply_nrt_default(m->tab.nl); }

// doubles of the k_apply_v coefficient layout for the whole mesh
static size_t vcoef_size(const f25_mesh *m, int *nrt, int *rt, int *cw) {
    const int nl = m->tab.nl, t = apply_nrt(m);
    if (!apply_nrt_ok(nl, t)) return 0;
    *nrt = t;
    *rt = (m->tab.nen + t - 1) / t;
    *cw = (5 * *rt + 31) / 32 * 32;
    return (size_t)m->d.nelem * 3 * *nrt * m->tab.nen * *cw;
}

// element blocks (NBLK per element) -> k_apply_v coefficient layout; the
// buffer *vcf is (re)allocated to the layout's size
static int vcoef(f25_mesh *m, const double *blk, double **vcf) {
    int nrt = 0, rt = 0, cw = 0;
    const size_t n = vcoef_size(m, &nrt, &rt, &cw);
    if (!n) FAIL(F25_ERR_ARG, "unsupported order");
    size_t *cap = (vcf == &m->d_vcf) ? &m->vcf_cap : &m->dvcf_cap;
    if (!*vcf || *cap < n) {
        dfree(*vcf);
        *cap = 0;
        TRY(dalloc(vcf, n));
        *cap = n;
    }
    const int nb = (int)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_vcoef, dim3(nb), dim3(256), 0, m->st, m->tab.nen, nrt, rt, cw, m->d.nelem, blk, *vcf);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int f25_set_apply_tiles(f25_mesh *m, int nrt) {
    if (!m) FAIL(F25_ERR_ARG, "null handle");
    if (nrt != 0 && !apply_nrt_ok(m->tab.nl, nrt)) FAIL(F25_ERR_ARG, "row tile count not compiled for this order");
    HIPCHK(hipSetDevice(m->dev));
    if (m->ag.exec || m->br.exec) {   // a captured product holds the previous kernel
        HIPCHK(hipStreamSynchronize(m->st));
        if (m->ag.exec) (void)hipGraphExecDestroy(m->ag.exec);
        if (m->br.exec) (void)hipGraphExecDestroy(m->br.exec);
        m->ag.exec = nullptr;
        m->br.exec = nullptr;
    }
    const int old = m->apply_nrt;
    m->apply_nrt = nrt;
    if (m->have_model && apply_nrt(m) != (old ? old : apply_nrt_default(m->tab.nl))) {
        const int r = vcoef(m, m->d_blk, &m->d_vcf);
        if (r) return r;
        HIPCHK(hipStreamSynchronize(m->st));
    }
    return 0;
}

// zero rows and store sink of the FP64 product, npair values each
static int ensure_vpad(f25_mesh *m, int npair) {
    if (npair > m->vpad_cap) {
        dfree(m->d_vpad);
        m->vpad_cap = 0;
        TRY(dalloc(&m->d_vpad, 4 * (size_t)npair));
        fill_async(m->d_vpad, 2 * (size_t)npair, 0.0, m->st);
        m->vpad_cap = npair;
    }
    return 0;
}

// the launches of the FP64 product (ensure_vpad first)
static int launch_f64(f25_mesh *m, const double *om, const double *kv, int npair, const double2 *x, double2 *y,
                      const double *vcf) {
    switch (m->tab.nl * 8 + apply_nrt(m)) {
#define F25_AV(N, T) case N * 8 + T: return launch_apply_v<N, T>(m, om, kv, npair, x, y, vcf);
        F25_AV(3, 1) F25_AV(3, 2) F25_AV(4, 1) F25_AV(4, 2) F25_AV(5, 1) F25_AV(5, 2) F25_AV(5, 3)
        F25_AV(6, 2) F25_AV(6, 3) F25_AV(6, 4) F25_AV(7, 2) F25_AV(7, 3) F25_AV(7, 4)
#undef F25_AV
    }
    FAIL(F25_ERR_ARG, "unsupported order");
}

static int apply_f64(f25_mesh *m, const double *om, const double *kv, int npair, const double2 *x, double2 *y,
                     const double *vcf) {
    TRY(ensure_vpad(m, npair));
    return launch_f64(m, om, kv, npair, x, y, vcf);
}

// ---- elements of the batched product integrated per omega ---------------
// The PML elements, or every element in a dispersive medium (its moduli
// depend on the frequency through VCMPLX): their colour order and colour
// offsets.
static const std::vector<int> &omega_cptr(const f25_mesh *m) { return m->disp ? m->cptr : m->pml_cptr; }
static const int *omega_order(const f25_mesh *m) { return m->disp ? m->d_order : m->d_pml_order; }
static size_t omega_nelem(const f25_mesh *m) { return m->disp ? m->order.size() : m->pml_order.size(); }

// Distinct omegas of the batch (first-appearance order), the pairs of each
// in groups of at most 16, and the element matrices for those omegas
// (rebuilt when the omegas, the model or the medium change).  No-op without
// such elements.  h_om: the batch's omegas on the host.
static int pml_prepare(f25_mesh *m, int npair, const double *h_om) {
    auto &P = m->pa;
    P.ngrp = 0;
    if (!omega_nelem(m)) return 0;
    std::vector<double> uo;
    std::vector<int> fi(npair);
    for (int b = 0; b < npair; ++b) {
        size_t k = 0;
        while (k < uo.size() && uo[k] != h_om[b]) ++k;
        if (k == uo.size()) uo.push_back(h_om[b]);
        fi[b] = (int)k;
    }
    P.grp.clear();
    P.gpair.clear();
    for (size_t f = 0; f < uo.size(); ++f) {
        const int first = (int)P.gpair.size();
        for (int b = 0; b < npair; ++b)
            if (fi[b] == (int)f) P.gpair.push_back(b);
        for (int g0 = first; g0 < (int)P.gpair.size(); g0 += 16) {
            P.grp.push_back((int)f);
            P.grp.push_back(g0);
            P.grp.push_back(std::min(16, (int)P.gpair.size() - g0));
        }
    }
    P.ngrp = (int)P.grp.size() / 3;
    if ((int)P.grp.size() > P.grp_cap) {
        dfree(P.d_grp);
        P.grp_cap = 0;
        TRY(dalloc(&P.d_grp, P.grp.size()));
        P.grp_cap = (int)P.grp.size();
    }
    if ((int)P.gpair.size() > P.gpair_cap) {
        dfree(P.d_gpair);
        P.gpair_cap = 0;
        TRY(dalloc(&P.d_gpair, P.gpair.size()));
        P.gpair_cap = (int)P.gpair.size();
    }
    HIPCHK(hipMemcpyAsync(P.d_grp, P.grp.data(), P.grp.size() * sizeof(int), hipMemcpyHostToDevice, m->st));
    HIPCHK(hipMemcpyAsync(P.d_gpair, P.gpair.data(), P.gpair.size() * sizeof(int), hipMemcpyHostToDevice, m->st));
    if (P.valid && P.disp == m->disp && uo == P.om) return 0;
    const int nen = m->tab.nen, nee = 3 * nen, nom = (int)uo.size();
    const size_t nm = omega_nelem(m) * (size_t)nom * 3 * nee * nee;
    if (nm > P.mats_cap) {
        dfree(P.d_mats);
        P.mats_cap = 0;
        TRY(dalloc(&P.d_mats, nm));
        P.mats_cap = nm;
    }
    if (nom > P.om_cap) {
        dfree(P.d_om);
        P.om_cap = 0;
        TRY(dalloc(&P.d_om, (size_t)nom));
        P.om_cap = nom;
    }
    HIPCHK(hipMemcpyAsync(P.d_om, uo.data(), nom * sizeof(double), hipMemcpyHostToDevice, m->st));
    PmlPar p{};
    p.xwidth = m->d.xwidth;
    p.zwidth = m->d.zwidth;
    p.rcoeff = m->d.pml[0]; p.dmax = m->d.pml[1]; p.kappa = m->d.pml[2]; p.fcent = m->d.pml[3];
    p.freq0 = m->freq0;
    const size_t lds = (8 * 9 + 3 * 8 * nen) * sizeof(double);
    const std::vector<int> &cp = omega_cptr(m);
    for (size_t c = 0; c + 1 < cp.size(); ++c) {
        const int n = cp[c + 1] - cp[c];
        if (!n) continue;
        hipLaunchKernelGGL(k_pml_split, dim3(n, nom), dim3(256), lds, m->st, nen, m->tab.nq1, cp[c], nom,
                           omega_order(m) + cp[c], P.d_om, m->d_shl, m->d_wq, m->d_xq, m->d_ieng, m->d_xl, m->d_zl,
                           m->d_xd, m->d_zd, m->d_cdom, m->d_rho, m->d_lam, m->d_mu, m->disp ? m->d_qp : nullptr,
                           m->disp ? m->d_qs : nullptr, p, P.d_mats, m->d_err);
    }
    HIPCHK(hipGetLastError());
    TRY(check_err(m));
    P.om = uo;
    P.disp = m->disp;
    P.valid = true;
    return 0;
}

// y += S_e x over the elements held per omega, colour by colour (after the
// product, or after a zero fill of y in a dispersive medium)
template <typename XT>
static int pml_launch(f25_mesh *m, int npair, const double *d_ky, const XT *x, XT *y) {
    auto &P = m->pa;
    if (!P.ngrp) return 0;
    const int nen = m->tab.nen, nom = (int)P.om.size();
    const size_t lds = (size_t)3 * nen * 16 * sizeof(double2);
    const std::vector<int> &cp = omega_cptr(m);
    for (size_t c = 0; c + 1 < cp.size(); ++c) {
        const int n = cp[c + 1] - cp[c];
        if (!n) continue;
        hipLaunchKernelGGL(k_apply_pml<XT>, dim3(n, P.ngrp), dim3(256), lds, m->st, nen, cp[c], nom,
                           omega_order(m) + cp[c], m->d_lm, P.d_grp, P.d_gpair, npair, d_ky, P.d_mats, x, y);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

// The whole operator product in a dispersive medium: y = 0, then every
// element's per-omega matrix
template <typename XT>
static int disp_product(f25_mesh *m, int npair, const double *d_ky, const XT *x, XT *y) {
    fill_async(y, (size_t)m->g.ndof * npair, XT{0, 0}, m->st);
    return pml_launch(m, npair, d_ky, x, y);
}

// F25_APPLY_GRAPH: capture the launches once per argument set, then replay
static int apply_f64_graph(f25_mesh *m, const double *om, const double *kv, int npair, const double2 *x,
                           double2 *y, const double *vcf) {
    TRY(ensure_vpad(m, npair));
    auto &G = m->ag;
    const auto &P = m->pa;
    const int nom = (int)P.om.size();
    if (!G.exec || G.x != x || G.y != y || G.npair != npair || G.vcf != vcf || G.om != om || G.kv != kv ||
        G.pad != m->d_vpad || G.mats != P.d_mats || G.grp != P.d_grp || G.gpair != P.d_gpair || G.nom != nom ||
        G.ngrp != P.ngrp || G.disp != m->disp) {
        if (G.exec) {
            HIPCHK(hipStreamSynchronize(m->st));
            (void)hipGraphExecDestroy(G.exec);
            G.exec = nullptr;
        }
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(m->st, hipStreamCaptureModeThreadLocal));
        int r = m->disp ? disp_product(m, npair, kv, x, y) : launch_f64(m, om, kv, npair, x, y, vcf);
        if (!r && !m->disp) r = pml_launch(m, npair, kv, x, y);
        hipError_t e = hipStreamEndCapture(m->st, &g);
        if (r) {
            if (g) (void)hipGraphDestroy(g);
            return r;
        }
        if (e != hipSuccess) FAIL(F25_ERR_HIP, hipGetErrorString(e));
        e = hipGraphInstantiate(&G.exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            G.exec = nullptr;
            FAIL(F25_ERR_HIP, hipGetErrorString(e));
        }
        G.x = x; G.y = y; G.npair = npair; G.vcf = vcf; G.om = om; G.kv = kv; G.pad = m->d_vpad;
        G.mats = P.d_mats; G.grp = P.d_grp; G.gpair = P.d_gpair; G.nom = nom; G.ngrp = P.ngrp; G.disp = m->disp;
    }
    HIPCHK(hipGraphLaunch(G.exec, m->st));
    return 0;
}

// h_om: the batch's omegas on the host when the PML elements are to be
// included (the operator product), null for products on derivative blocks
// (the PML anchors' parameters are held fixed there).
static int apply_core(f25_mesh *m, int npair, const double *d_om, const double *d_ky, int mode,
                      const void *x, void *y, const double *vcf, const double *h_om) {
    // every DOF is written by the first element touching it: no zero fill
    const bool graph = (mode & F25_APPLY_GRAPH) != 0;
    mode &= ~F25_APPLY_GRAPH;
    if (mode != F25_APPLY_F64 && mode != F25_APPLY_F32 && mode != F25_APPLY_F32_REFINED)
        FAIL(F25_ERR_ARG, "bad apply mode");
    if (graph && mode != F25_APPLY_F64) FAIL(F25_ERR_ARG, "graph replay is for the FP64 product");
    if (h_om) TRY(pml_prepare(m, npair, h_om));
    else m->pa.ngrp = 0;
    if (m->disp && !h_om) FAIL(F25_ERR_STATE, "products on element blocks need an elastic medium");
    HIPCHK(hipEventRecord(m->ev0, m->st));
    if (m->disp && !graph) {
        // every element per omega: no element blocks
        if (mode == F25_APPLY_F32) TRY(disp_product(m, npair, d_ky, (const float2 *)x, (float2 *)y));
        else TRY(disp_product(m, npair, d_ky, (const double2 *)x, (double2 *)y));
    } else if (mode == F25_APPLY_F64 && graph)
        TRY(apply_f64_graph(m, d_om, d_ky, npair, (const double2 *)x, (double2 *)y, vcf));
    else if (mode == F25_APPLY_F64) {
        TRY(apply_f64(m, d_om, d_ky, npair, (const double2 *)x, (double2 *)y, vcf));
        TRY(pml_launch(m, npair, d_ky, (const double2 *)x, (double2 *)y));
    } else if (mode == F25_APPLY_F32) {
        TRY((dispatch_apply<float, float2, false>(m, d_om, d_ky, npair, (const float2 *)x, (float2 *)y,
                                                  m->d_blk32, nullptr)));
        TRY(pml_launch(m, npair, d_ky, (const float2 *)x, (float2 *)y));
    } else {
        if (npair > m->colmax_cap) {
            dfree(m->d_colmax);
            TRY(dalloc(&m->d_colmax, (size_t)npair));
            m->colmax_cap = npair;
        }
        fill_async(m->d_colmax, (size_t)npair, 0ull, m->st);
        const int gy = (int)std::min<int64_t>(1024, std::max<int64_t>(1, m->g.ndof / 64));
        hipLaunchKernelGGL(k_colmax, dim3((npair + 63) / 64, gy), dim3(256), 0, m->st, (int64_t)m->g.ndof,
                           npair, (const double2 *)x, m->d_colmax);
        HIPCHK(hipGetLastError());
        TRY((dispatch_apply<float, double2, true>(m, d_om, d_ky, npair, (const double2 *)x, (double2 *)y,
                                                  m->d_blk_hi, m->d_blk_lo)));
        TRY(pml_launch(m, npair, d_ky, (const double2 *)x, (double2 *)y));
    }
    HIPCHK(hipEventRecord(m->ev1, m->st));
    HIPCHK(hipStreamSynchronize(m->st));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, m->ev0, m->ev1));
    record_ms("apply", ms);
    return 0;
}

static int stage_omky(f25_mesh *m, int n, const double *omega, const double *ky) {
    if (n > m->omky_cap) {
        dfree(m->d_omky);
        TRY(dalloc(&m->d_omky, 2 * (size_t)n));
        m->omky_cap = n;
    }
    HIPCHK(hipMemcpyAsync(m->d_omky, omega, n * sizeof(double), hipMemcpyHostToDevice, m->st));
    HIPCHK(hipMemcpyAsync(m->d_omky + m->omky_cap, ky, n * sizeof(double), hipMemcpyHostToDevice, m->st));
    return 0;
}

extern "C" int f25_apply_batch_dev(f25_mesh *m, int npair, const double *omega, const double *ky,
                                   int mode, const void *x, void *y) {
    if (!m || npair <= 0 || !omega || !ky || !x || !y) FAIL(F25_ERR_ARG, "invalid apply arguments");
    if (!m->have_model) FAIL(F25_ERR_STATE, "f25_set_model has not been called");
    HIPCHK(hipSetDevice(m->dev));
    TRY(stage_omky(m, npair, omega, ky));
    return apply_core(m, npair, m->d_omky, m->d_omky + m->omky_cap, mode, x, y, m->d_vcf, omega);
}

extern "C" int f25_residual_batch_dev(f25_mesh *m, int npair, const double *omega, const double *ky,
                                      const void *u, const void *f, void *r, double *rnorm) {
    if (!m || npair <= 0 || !omega || !ky || !u || !f || !r || !rnorm || u == r || f == r)
        FAIL(F25_ERR_ARG, "invalid residual arguments");
    if (!m->have_model) FAIL(F25_ERR_STATE, "f25_set_model has not been called");
    HIPCHK(hipSetDevice(m->dev));
    TRY(stage_omky(m, npair, omega, ky));
    TRY(apply_core(m, npair, m->d_omky, m->d_omky + m->omky_cap, F25_APPLY_F64, u, r, m->d_vcf, omega));
    const int gy = (int)std::min<int64_t>(256, std::max<int64_t>(1, m->g.ndof / 256));
    double *part = nullptr, *sums = nullptr;
    int rc = dalloc(&part, (size_t)gy * npair * 2);
    if (!rc) rc = dalloc(&sums, (size_t)npair * 2);
    if (!rc) {
        hipLaunchKernelGGL(k_resid, dim3((npair + 63) / 64, gy), dim3(256), 0, m->st, (int64_t)m->g.ndof, npair,
                           (const double2 *)f, (double2 *)r, part);
        hipLaunchKernelGGL(k_resid_sum, dim3((2 * npair + 255) / 256), dim3(256), 0, m->st, gy, npair, part, sums);
        std::vector<double> h(2 * (size_t)npair);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(h.data(), sums, h.size() * sizeof(double), hipMemcpyDeviceToHost, m->st);
        if (e == hipSuccess) e = hipStreamSynchronize(m->st);
        if (e != hipSuccess) { g_err = hipGetErrorString(e); rc = F25_ERR_HIP; }
        for (int b = 0; b < npair && !rc; ++b)
            rnorm[b] = h[2 * b + 1] > 0.0 ? std::sqrt(h[2 * b] / h[2 * b + 1]) : std::sqrt(h[2 * b]);
    }
    dfree(part);
    dfree(sums);
    return rc;
}

template <typename T>
static int transpose_dev(hipStream_t s, const T *in, T *out, int rows, int cols) {
    const size_t nt = (size_t)((cols + 15) / 16) * ((rows + 15) / 16);
    if (nt > 0x7fffffff) FAIL(F25_ERR_ARG, "transpose too large");
    hipLaunchKernelGGL(k_transpose_c<T>, dim3((unsigned)nt), dim3(16, 16), 0, s, in, out, rows, cols);
    HIPCHK(hipGetLastError());
    return 0;
}

// ---- Bielak sources in the batch ------------------------------------------
// The ring: elements with a 'B' and an 'E' DOF (only they couple the two
// classes, asmble.f90:356-363), coloured; pass 0 reads the 'E' DOFs and
// writes the 'B' rows, pass 1 reads 'B' and writes 'E' (the DOF maps hold -1
// elsewhere: read zero, store to the sink).
static int bielak_prepare(f25_mesh *m) {
    auto &B = m->br;
    if (B.built) return 0;
    const int nelem = m->d.nelem, nen = m->tab.nen, nee = 3 * nen;
    const std::vector<char> &cnp = m->g.cnp;
    std::vector<int> ring, sub;
    for (int e = 0; e < nelem; ++e) {
        bool hb = false, he = false;
        for (int j = 0; j < nee; ++j) {
            const int d = m->g.lm[(size_t)e * nee + j];
            if (d <= 0) continue;
            hb |= cnp[d - 1] == 'B';
            he |= cnp[d - 1] == 'E';
        }
        if (!(hb && he)) continue;
        if (m->cdom[e] == 'A') FAIL(F25_ERR_ARG, "a Bielak ring element lies in the PML");
        ring.push_back(e);
        sub.insert(sub.end(), m->ieng.begin() + 4 * e, m->ieng.begin() + 4 * e + 4);
    }
    std::vector<int> order, ptr;
    if (!ring.empty()) colour_elements((int)ring.size(), m->d.nnpg, sub.data(), order, ptr);
    else ptr.assign(1, 0);
    std::vector<int> ord(ring.size());
    for (size_t i = 0; i < ring.size(); ++i) ord[i] = ring[order[i]];
    const char cls[4] = {'E', 'B', 'B', 'E'};
    std::vector<int> lmv[4];
    for (int k = 0; k < 4; ++k) lmv[k].assign((size_t)nelem * nee, -1);
    std::vector<char> isb(m->g.ndof, 0);
    for (int e : ring)
        for (int j = 0; j < nee; ++j) {
            const int d = m->g.lm[(size_t)e * nee + j];
            if (d <= 0) continue;
            for (int k = 0; k < 4; ++k)
                if (cnp[d - 1] == cls[k]) lmv[k][(size_t)e * nee + j] = d - 1;
            if (cnp[d - 1] == 'B') isb[d - 1] = 1;
        }
    std::vector<int> brows;
    for (int d = 0; d < m->g.ndof; ++d)
        if (isb[d]) brows.push_back(d);
    TRY(upload(&B.d_order, ord.data(), ord.size(), m->st));
    for (int k = 0; k < 4; ++k) TRY(upload(&B.d_lm[k], lmv[k].data(), lmv[k].size(), m->st));
    TRY(dalloc(&B.d_first0, (size_t)nelem));
    fill_async(B.d_first0, (size_t)nelem, 0ull, m->st);
    TRY(upload(&B.d_brows, brows.data(), brows.size(), m->st));
    HIPCHK(hipStreamSynchronize(m->st));
    B.nb = (int)brows.size();
    B.cptr = ptr;
    B.built = true;
    return 0;
}

template <int NL, int NRT>
static int launch_ring(f25_mesh *m, int pass, const double *om, const double *kv, int npair, const double2 *x,
                       double2 *y) {
    auto &B = m->br;
    const double2 *zpad = reinterpret_cast<const double2 *>(m->d_vpad);
    double2 *sink = reinterpret_cast<double2 *>(m->d_vpad) + m->vpad_cap;
    for (size_t c = 0; c + 1 < B.cptr.size(); ++c) {
        const int n = B.cptr[c + 1] - B.cptr[c];
        if (!n) continue;
        const size_t nwg = apply_v_grid<NL, NRT>(n, npair);
        if (nwg > 0x7fffffff) FAIL(F25_ERR_ARG, "too many pairs for one launch");
        hipLaunchKernelGGL((k_apply_v<NL, NRT>), dim3((unsigned)nwg), dim3(VShape<NL, NRT>::NT), 0, m->st,
                           B.d_order + B.cptr[c], n, m->d_vcf, B.d_lm[2 * pass], B.d_lm[2 * pass + 1], B.d_first0,
                           npair, om, kv, x, y, zpad, sink);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

static int ring_pass(f25_mesh *m, int pass, const double *om, const double *kv, int npair, const double2 *x,
                     double2 *y) {
    switch (m->tab.nl * 8 + apply_nrt(m)) {
#define F25_RV(N, T) case N * 8 + T: return launch_ring<N, T>(m, pass, om, kv, npair, x, y);
        F25_RV(3, 1) F25_RV(3, 2) F25_RV(4, 1) F25_RV(4, 2) F25_RV(5, 1) F25_RV(5, 2) F25_RV(5, 3)
        F25_RV(6, 2) F25_RV(6, 3) F25_RV(6, 4) F25_RV(7, 2) F25_RV(7, 3) F25_RV(7, 4)
#undef F25_RV
    }
    FAIL(F25_ERR_ARG, "unsupported order");
}

// the effective force's launches (bielak_prepare and ensure_vpad first):
// zero fill, pass 0, 'B' rows negated, pass 1; no host synchronisation
static int peff_launches(f25_mesh *m, int npair, const double *om, const double *kv, const double2 *u,
                         double2 *y) {
    auto &B = m->br;
    fill_async(y, (size_t)m->g.ndof * npair, make_double2(0.0, 0.0), m->st);
    TRY(ring_pass(m, 0, om, kv, npair, u, y));
    if (B.nb) {
        const int64_t tot = (int64_t)B.nb * npair;
        const int nb = (int)std::min<int64_t>((tot + 255) / 256, 65536);
        hipLaunchKernelGGL(k_negate_rows, dim3(nb), dim3(256), 0, m->st, B.nb, npair, B.d_brows, y);
    }
    TRY(ring_pass(m, 1, om, kv, npair, u, y));
    HIPCHK(hipGetLastError());
    return 0;
}

// ASMB_PEFF (asmble.f90:318-377) for npair (omega, ky) pairs at once, FP64:
// peff = 0 outside the 'B' and 'E' rows, -S_BE u_E on 'B' rows, +S_EB u_E on
// 'E' rows, S(omega, ky) from the element blocks.
extern "C" int f25_peff_batch_dev(f25_mesh *m, int npair, const double *omega, const double *ky,
                                  const void *ue_dev, void *peff_dev) {
    if (!m || npair <= 0 || !omega || !ky || !ue_dev || !peff_dev || ue_dev == peff_dev)
        FAIL(F25_ERR_ARG, "invalid arguments");
    if (!m->have_model) FAIL(F25_ERR_STATE, "f25_set_model has not been called");
    if (m->disp) FAIL(F25_ERR_ARG, "dispersive medium: assemble per pair and use f25_asmb_peff");
    HIPCHK(hipSetDevice(m->dev));
    TRY(bielak_prepare(m));
    TRY(stage_omky(m, npair, omega, ky));
    TRY(ensure_vpad(m, npair));
    const double *om = m->d_omky, *kv = m->d_omky + m->omky_cap;
    HIPCHK(hipEventRecord(m->ev0, m->st));
    TRY(peff_launches(m, npair, om, kv, (const double2 *)ue_dev, (double2 *)peff_dev));
    HIPCHK(hipEventRecord(m->ev1, m->st));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(m->st));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, m->ev0, m->ev1) == hipSuccess) record_ms("peff_batch", ms);
    return 0;
}

// FILL_UWAVE_QMESH's DOF positions: each DOF at the GLL point of the first
// element (element, eta, xi order) holding it, placed by CSF2DND
static int uwave_prepare(f25_mesh *m) {
    auto &B = m->br;
    if (B.uw) return 0;
    const int nl = m->tab.nl, nen = m->tab.nen, ndof = m->g.ndof;
    std::vector<double2> xz(ndof, make_double2(0.0, 0.0));
    std::vector<signed char> comp(ndof, 0);
    std::vector<char> init(ndof, 0);
    for (int e = 0; e < m->d.nelem; ++e)
        for (int ie = 0; ie < nl; ++ie)
            for (int ix = 0; ix < nl; ++ix) {
                const int *lm = &m->g.lm[((size_t)e * nen + ie * nl + ix) * 3];
                bool seen = false;
                for (int i = 0; i < 3; ++i) seen |= lm[i] > 0 && init[lm[i] - 1];
                if (seen) continue;
                double x, z;
                gll_xz(m->d, e, m->tab.xipts[ix], m->tab.xipts[ie], x, z);
                for (int i = 0; i < 3; ++i)
                    if (lm[i] > 0) {
                        init[lm[i] - 1] = 1;
                        xz[lm[i] - 1] = make_double2(x, z);
                        comp[lm[i] - 1] = (signed char)i;
                    }
            }
    for (int d = 0; d < ndof; ++d)
        if (!init[d]) FAIL(F25_ERR_GRAPH, "a DOF belongs to no element");
    TRY(upload(&B.d_xz, xz.data(), xz.size(), m->st));
    TRY(upload(&B.d_comp, comp.data(), comp.size(), m->st));
    HIPCHK(hipStreamSynchronize(m->st));
    B.uw = true;
    return 0;
}

// the plane waves' parameters (kx, kz, three polarisation factors) staged
// on the device; the factors as FILL_UWAVE_QMESH forms them in FP64
static int uwave_params(f25_mesh *m, int nwave, const double *kappax, const double *kappaz, const double *baz,
                        const double *aoi) {
    auto &B = m->br;
    const double pi180 = 0.017453292519943295;
    std::vector<double> pw(5 * (size_t)nwave);
    for (int w = 0; w < nwave; ++w) {
        const double cosaz = std::cos(baz[w] * pi180), sinaz = std::sin(baz[w] * pi180);
        const double cosi = std::cos(aoi[w] * pi180), sini = std::sin(aoi[w] * pi180);
        pw[5 * w + 0] = kappax[w];
        pw[5 * w + 1] = kappaz[w];
        pw[5 * w + 2] = cosaz * sini;
        pw[5 * w + 3] = sinaz * sini;
        pw[5 * w + 4] = cosi;
    }
    if (5 * nwave > B.pw_cap) {
        dfree(B.d_pw);
        B.pw_cap = 0;
        TRY(dalloc(&B.d_pw, 5 * (size_t)nwave));
        B.pw_cap = 5 * nwave;
    }
    HIPCHK(hipMemcpyAsync(B.d_pw, pw.data(), pw.size() * sizeof(double), hipMemcpyHostToDevice, m->st));
    return 0;
}

static int uwave_launch(f25_mesh *m, int nwave, double2 *ue) {
    auto &B = m->br;
    const int64_t tot = (int64_t)m->g.ndof * nwave;
    const int nb = (int)std::min<int64_t>((tot + 255) / 256, 1 << 20);
    hipLaunchKernelGGL(k_uwave, dim3(nb), dim3(256), 0, m->st, (int64_t)m->g.ndof, nwave, B.d_xz, B.d_comp, B.d_pw,
                       ue);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int f25_fill_uwave_dev(f25_mesh *m, int nwave, const double *kappax, const double *kappaz,
                                  const double *baz, const double *aoi, void *ue_dev) {
    if (!m || nwave <= 0 || !kappax || !kappaz || !baz || !aoi || !ue_dev) FAIL(F25_ERR_ARG, "invalid arguments");
    HIPCHK(hipSetDevice(m->dev));
    TRY(uwave_prepare(m));
    TRY(uwave_params(m, nwave, kappax, kappaz, baz, aoi));
    HIPCHK(hipEventRecord(m->ev0, m->st));
    TRY(uwave_launch(m, nwave, (double2 *)ue_dev));
    HIPCHK(hipEventRecord(m->ev1, m->st));
    HIPCHK(hipStreamSynchronize(m->st));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, m->ev0, m->ev1) == hipSuccess) record_ms("uwave", ms);
    return 0;
}

// Plane-wave sources of a batch: pair b's wavefield u_E (FILL_UWAVE_QMESH
// with kappax[b], kappaz[b], baz[b], aoi[b]) and its effective force at
// (omega[b], ky[b]).  flags & F25_APPLY_GRAPH: the launches are captured in
// a HIP graph on the first call and replayed while npair and the buffers
// stay the same (the parameters are staged before each replay).
extern "C" int f25_bielak_rhs_dev(f25_mesh *m, int npair, const double *omega, const double *ky,
                                  const double *kappax, const double *kappaz, const double *baz,
                                  const double *aoi, void *ue_dev, void *peff_dev, int flags) {
    if (!m || npair <= 0 || !omega || !ky || !kappax || !kappaz || !baz || !aoi || !ue_dev || !peff_dev ||
        ue_dev == peff_dev || (flags & ~F25_APPLY_GRAPH))
        FAIL(F25_ERR_ARG, "invalid arguments");
    if (!m->have_model) FAIL(F25_ERR_STATE, "f25_set_model has not been called");
    if (m->disp) FAIL(F25_ERR_ARG, "dispersive medium: assemble per pair and use f25_asmb_peff");
    HIPCHK(hipSetDevice(m->dev));
    TRY(bielak_prepare(m));
    TRY(uwave_prepare(m));
    auto &B = m->br;
    if (B.exec && (B.npair != npair || B.ue != ue_dev || B.peff != peff_dev)) {
        HIPCHK(hipStreamSynchronize(m->st));
        (void)hipGraphExecDestroy(B.exec);
        B.exec = nullptr;
    }
    TRY(stage_omky(m, npair, omega, ky));
    TRY(ensure_vpad(m, npair));
    TRY(uwave_params(m, npair, kappax, kappaz, baz, aoi));
    const double *om = m->d_omky, *kv = m->d_omky + m->omky_cap;
    // staging buffers may have moved since the capture
    if (B.exec && (B.om != om || B.pw != B.d_pw || B.pad != m->d_vpad || B.vcf != m->d_vcf)) {
        HIPCHK(hipStreamSynchronize(m->st));
        (void)hipGraphExecDestroy(B.exec);
        B.exec = nullptr;
    }
    HIPCHK(hipEventRecord(m->ev0, m->st));
    if (!(flags & F25_APPLY_GRAPH)) {
        TRY(uwave_launch(m, npair, (double2 *)ue_dev));
        TRY(peff_launches(m, npair, om, kv, (const double2 *)ue_dev, (double2 *)peff_dev));
    } else {
        if (!B.exec) {
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(m->st, hipStreamCaptureModeThreadLocal));
            int r = uwave_launch(m, npair, (double2 *)ue_dev);
            if (!r) r = peff_launches(m, npair, om, kv, (const double2 *)ue_dev, (double2 *)peff_dev);
            hipError_t e = hipStreamEndCapture(m->st, &g);
            if (r) {
                if (g) (void)hipGraphDestroy(g);
                return r;
            }
            if (e != hipSuccess) FAIL(F25_ERR_HIP, hipGetErrorString(e));
            e = hipGraphInstantiate(&B.exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                B.exec = nullptr;
                FAIL(F25_ERR_HIP, hipGetErrorString(e));
            }
            B.ue = ue_dev; B.peff = peff_dev; B.npair = npair; B.om = om; B.pw = B.d_pw; B.pad = m->d_vpad;
            B.vcf = m->d_vcf;
        }
        HIPCHK(hipGraphLaunch(B.exec, m->st));
    }
    HIPCHK(hipEventRecord(m->ev1, m->st));
    HIPCHK(hipStreamSynchronize(m->st));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, m->ev0, m->ev1) == hipSuccess) record_ms("bielak_rhs", ms);
    return 0;
}

extern "C" int f25_apply_batch(f25_mesh *m, int npair, const double *omega, const double *ky, int mode,
                               const void *x, void *y) {
    if (!m || npair <= 0 || !x || !y) FAIL(F25_ERR_ARG, "invalid apply arguments");
    if (mode == F25_APPLY_F32) FAIL(F25_ERR_ARG, "host wrapper takes double complex data; use F64 or F32_REFINED");
    HIPCHK(hipSetDevice(m->dev));
    const size_t n = (size_t)m->g.ndof * npair;
    double2 *a = nullptr, *b = nullptr;
    TRY(dalloc(&a, n));
    TRY(dalloc(&b, n));
    int r = 0;
    hipError_t e = hipMemcpyAsync(a, x, n * 16, hipMemcpyHostToDevice, m->st);
    if (e != hipSuccess) r = F25_ERR_HIP;
    if (!r) r = transpose_dev(m->st, a, b, npair, m->g.ndof);
    if (!r) r = f25_apply_batch_dev(m, npair, omega, ky, mode, b, a);
    if (!r) r = transpose_dev(m->st, a, b, m->g.ndof, npair);
    if (!r) {
        e = hipMemcpyAsync(y, b, n * 16, hipMemcpyDeviceToHost, m->st);
        if (e == hipSuccess) e = hipStreamSynchronize(m->st);
        if (e != hipSuccess) { g_err = hipGetErrorString(e); r = F25_ERR_HIP; }
    }
    dfree(a);
    dfree(b);
    return r;
}

// ---------------------------------------------------------------------------
// quadrature points per lane of k_grad_sf for the order
template <typename R>
static int grad_ppl(int nl) {
    return nl == 7 ? GradOcc<7, R>::PPL : 1;
}

// threads of a k_grad_sf workgroup (its launch shape for the order)
template <typename R>
static int grad_threads(int nl) {
    switch (nl) {
        case 3: return GradOcc<3, R>::NT;
        case 4: return GradOcc<4, R>::NT;
        case 5: return GradOcc<5, R>::NT;
        case 6: return GradOcc<6, R>::NT;
        default: return GradOcc<7, R>::NT;
    }
}

template <typename R>
static int grad_launch(f25_mesh *m, int ntrip, const double *d_om, const double *d_ky, const void *u, const void *l) {
    typedef typename CxOf<R>::T C2;
    const int nen = m->tab.nen;
    // every order up to p6 runs the sum-factorised kernel
    typedef void (*GradSf)(const double *, const double *, const double *, const int *, const double *,
                           const double *, const char *, const int *, int, int64_t, const double *,
                           const double *, const C2 *, const C2 *, double *, int *);
    // trips per pass (GradOcc::T): with two, four workgroups fit per CU at
    // p4 (LDS and registers; 200x100, 128 triples: 9.4 ms)
    GradSf ksf = nullptr;
    int TS = 0;
    switch (m->tab.nl) {
        case 3: ksf = k_grad_sf<3, GradOcc<3, R>::T, R>; TS = GradOcc<3, R>::T; break;
        case 4: ksf = k_grad_sf<4, GradOcc<4, R>::T, R>; TS = GradOcc<4, R>::T; break;
        case 5: ksf = k_grad_sf<5, GradOcc<5, R>::T, R>; TS = GradOcc<5, R>::T; break;
        case 6: ksf = k_grad_sf<6, GradOcc<6, R>::T, R>; TS = GradOcc<6, R>::T; break;
        case 7: ksf = k_grad_sf<7, GradOcc<7, R>::T, R>; TS = GradOcc<7, R>::T; break;
        default: break;
    }
    if (!ksf) FAIL(F25_ERR_ARG, "unsupported order");
    const int nl = m->tab.nl, q1 = m->tab.nq1;
    // (the FP32 form adds the inverse Jacobian at the points, 4 values each)
    // (the FP64 two-point form adds the interleaved 1-D table, 2 N1 Q1 doubles)
    const size_t lsf = (size_t)TS * 2 * 3 * nen * sizeof(C2) + (size_t)TS * 2 * 3 * q1 * nl * 2 * sizeof(C2) +
                       (size_t)2 * nl * q1 * sizeof(R) +
                       (std::is_same<R, float>::value ? (size_t)4 * q1 * q1 * sizeof(R) : 0) +
                       (grad_ppl<R>(nl) == 2 ? (size_t)2 * nl * q1 * sizeof(double) : 0);
    if (lsf > 64 * 1024)
        HIPCHK(hipFuncSetAttribute((const void *)ksf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lsf));
    HIPCHK(hipEventRecord(m->ev0, m->st));
    hipLaunchKernelGGL(ksf, dim3(m->d.nelem), dim3(grad_threads<R>(nl)), lsf, m->st, m->d_t1d, m->d_wq,
                       m->d_xq, m->d_ieng, m->d_xl, m->d_zl, m->d_cdom, m->d_lm, ntrip, (int64_t)m->g.ndof,
                       d_om, d_ky, (const C2 *)u, (const C2 *)l, m->d_part, m->d_err);
    return 0;
}

// fp32: complex64 fields, FP32 interpolation and densities per triple, FP64 sums
static int grad_core(f25_mesh *m, int ntrip, const double *d_om, const double *d_ky, const void *u,
                     const void *l, double *g, int chain, bool fp32 = false) {
    if (fp32) TRY(grad_launch<float>(m, ntrip, d_om, d_ky, u, l));
    else TRY(grad_launch<double>(m, ntrip, d_om, d_ky, u, l));
    HIPCHK(hipEventRecord(m->ev1, m->st));
    hipLaunchKernelGGL(k_grad_gather, dim3((m->d.nnpg + 255) / 256), dim3(256), 0, m->st, m->d.nnpg,
                       m->ncon_anchor, m->d_mconn, m->d_ieng, m->d_part, m->d_rho, m->d_lam, m->d_mu,
                       m->d_fixa, chain, g);
    HIPCHK(hipGetLastError());
    TRY(check_err(m));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, m->ev0, m->ev1));
    record_ms("grad", ms);
    return 0;
}

extern "C" int f25_gradient_dev(f25_mesh *m, int ntrip, const double *omega, const double *ky,
                                const void *u, const void *lam, double *g) {
    if (m) TRY(require_elastic(m));
    if (!m || ntrip <= 0 || !u || !lam || !g) FAIL(F25_ERR_ARG, "invalid gradient arguments");
    if (!m->have_model) FAIL(F25_ERR_STATE, "f25_set_model has not been called");
    HIPCHK(hipSetDevice(m->dev));
    TRY(stage_omky(m, ntrip, omega, ky));
    return grad_core(m, ntrip, m->d_omky, m->d_omky + m->omky_cap, u, lam, g, 1);
}

extern "C" int f25_gradient32_dev(f25_mesh *m, int ntrip, const double *omega, const double *ky,
                                  const void *u, const void *lam, double *g) {
    if (m) TRY(require_elastic(m));
    if (!m || ntrip <= 0 || !u || !lam || !g) FAIL(F25_ERR_ARG, "invalid gradient arguments");
    if (!m->have_model) FAIL(F25_ERR_STATE, "f25_set_model has not been called");
    HIPCHK(hipSetDevice(m->dev));
    TRY(stage_omky(m, ntrip, omega, ky));
    return grad_core(m, ntrip, m->d_omky, m->d_omky + m->omky_cap, u, lam, g, 1, true);
}

static int grad_host(f25_mesh *m, int ntrip, const double *omega, const double *ky, const void *u,
                     const void *lam, double *g, int chain) {
    if (m) TRY(require_elastic(m));
    if (!m || ntrip <= 0 || !u || !lam || !g) FAIL(F25_ERR_ARG, "invalid gradient arguments");
    if (!m->have_model) FAIL(F25_ERR_STATE, "f25_set_model has not been called");
    HIPCHK(hipSetDevice(m->dev));
    const size_t n = (size_t)m->g.ndof * ntrip;
    double2 *du = nullptr, *dl = nullptr;
    double *dg = nullptr, *d_om = nullptr, *d_ky = nullptr;
    TRY(upload(&du, (const double2 *)u, n, m->st));
    TRY(upload(&dl, (const double2 *)lam, n, m->st));
    TRY(upload(&dg, g, 3 * (size_t)m->d.nnpg, m->st));
    TRY(upload(&d_om, omega, ntrip, m->st));
    TRY(upload(&d_ky, ky, ntrip, m->st));
    int r = grad_core(m, ntrip, d_om, d_ky, du, dl, dg, chain);
    if (!r) {
        hipError_t e = hipMemcpyAsync(g, dg, 3 * (size_t)m->d.nnpg * 8, hipMemcpyDeviceToHost, m->st);
        if (e == hipSuccess) e = hipStreamSynchronize(m->st);
        if (e != hipSuccess) { g_err = hipGetErrorString(e); r = F25_ERR_HIP; }
    }
    dfree(du); dfree(dl); dfree(dg); dfree(d_om); dfree(d_ky);
    return r;
}

extern "C" int f25_gradient(f25_mesh *m, int ntrip, const double *omega, const double *ky, const void *u,
                            const void *lam, double *g) {
    return grad_host(m, ntrip, omega, ky, u, lam, g, 1);
}

// ---------------------------------------------------------------------------
static int inv_upload(f25_mesh *m) {
    m->have_inv = false;
    dfree(m->d_mygrad); dfree(m->d_jcsc); dfree(m->d_icsc); dfree(m->d_ew); dfree(m->d_dab_pos);
    TRY(upload(&m->d_mygrad, m->inv.mygrad.data(), m->inv.mygrad.size(), m->st));
    TRY(upload(&m->d_jcsc, m->inv.jcsc.data(), m->inv.jcsc.size(), m->st));
    TRY(upload(&m->d_icsc, m->inv.icsc.data(), m->inv.icsc.size(), m->st));
    TRY(upload(&m->d_ew, m->inv.elem_wts.data(), m->inv.elem_wts.size(), m->st));
    HIPCHK(hipStreamSynchronize(m->st));
    m->have_inv = true;
    return 0;
}

// CINVTYPE: PP (Vp), SS (Vs) with NVINV = 1, PS (both, NVINV = 2), as the
// driver maps them (gn25.f90:170-183)
static int inv_kind(const char *cinvtype, bool &lisp, bool &liss) {
    const char c0 = (char)toupper(cinvtype[0]), c1 = (char)toupper(cinvtype[1]);
    lisp = (c0 == 'P' && c1 == 'P') || (c0 == 'P' && c1 == 'S');
    liss = (c0 == 'S' && c1 == 'S') || (c0 == 'P' && c1 == 'S');
    if (!lisp && !liss) FAIL(F25_ERR_INVTYPE, "CINVTYPE must be PP, SS or PS (gn25.f90:170-183)");
    return 0;
}

// the variable of local column c: Vp for PP and for the first column of
// each PS pair (MYGRAD = 2 (INPINV - 1) + 1), Vs otherwise
static bool col_is_p(const InvInfo &inv, int c) {
    if (inv.nvinv == 2) return (inv.mygrad[c] - 1) % 2 == 0;
    return inv.lisp;
}

extern "C" int f25_inv_setup(f25_mesh *m, const char *cinvtype, double vpvs, int linrec, int lkbdry,
                             int nrec, const int *mrdof1, int sizes[6]) {
    if (!m || !cinvtype) FAIL(F25_ERR_ARG, "null argument");
    bool lisp, liss;
    TRY(inv_kind(cinvtype, lisp, liss));
    if (!linrec && nrec > 0 && !mrdof1) FAIL(F25_ERR_ARG, "receiver DOFs missing");
    HIPCHK(hipSetDevice(m->dev));
    InvInfo inv;
    int r = build_inv(m->d, m->g, lisp && liss ? 2 : 1, linrec, lkbdry, nrec, mrdof1, inv);
    if (r) FAIL(r, "inversion setup failed (no inversion points)");
    inv.lisp = lisp;
    inv.liss = liss;
    inv.vpvs = vpvs;
    m->inv = std::move(inv);
    if (sizes) {
        sizes[0] = m->inv.nnpinv; sizes[1] = m->inv.ncon; sizes[2] = m->inv.nnpgl;
        sizes[3] = m->inv.nz_fdist; sizes[4] = m->inv.mbufrhs; sizes[5] = m->inv.na35;
    }
    return inv_upload(m);
}

extern "C" int f25_inv_attach(f25_mesh *m, const char *cinvtype, double vpvs, int nvinv, int ncon,
                              const int *mconn, const int *maskg, int nnpgl, const int *mygrad,
                              const int *jcsc, int nz_fdist, const int *icsc, int na35,
                              const double *elem_wts) {
    if (!m || !cinvtype || !mconn || !maskg || !mygrad || !jcsc || !icsc || !elem_wts)
        FAIL(F25_ERR_ARG, "null argument");
    bool lisp, liss;
    TRY(inv_kind(cinvtype, lisp, liss));
    if (nvinv != (lisp && liss ? 2 : 1)) FAIL(F25_ERR_INVTYPE, "NVINV must be 1 for PP/SS and 2 for PS");
    const int nnpg = m->d.nnpg, nelem = m->d.nelem, ndof = m->g.ndof;
    if (ncon <= 0 || nnpgl <= 0 || nz_fdist < 0 || na35 <= 0) FAIL(F25_ERR_ARG, "bad sizes");
    if (jcsc[0] != 1 || jcsc[nnpgl] - 1 != nz_fdist) FAIL(F25_ERR_ARG, "JCSC_FDIST inconsistent with NZ_FDIST");
    InvInfo inv;
    inv.nvinv = nvinv;
    inv.ncon = ncon;
    inv.nnpgl = nnpgl;
    inv.nz_fdist = nz_fdist;
    inv.na35 = na35;
    inv.lisp = lisp;
    inv.liss = liss;
    inv.vpvs = vpvs;
    inv.mconn.assign(mconn, mconn + (size_t)ncon * nnpg);
    inv.maskg.assign(maskg, maskg + nnpg);
    inv.mygrad.assign(mygrad, mygrad + nnpgl);
    inv.jcsc.assign(jcsc, jcsc + nnpgl + 1);
    inv.icsc.assign(icsc, icsc + nz_fdist);
    inv.elem_wts.assign(elem_wts, elem_wts + nelem);
    // anchor of each column: (MASKG(INPG) - 1) NVINV + IVINV == MYGRAD(INPGL)
    // (hessloc.f90:66-72)
    const int ninv = na35 / nvinv;
    std::vector<int> anchor_of(std::max(ninv, 1) + 1, -1);
    for (int a = 0; a < nnpg; ++a) {
        const int v = maskg[a];
        if (v > 0 && v <= ninv && anchor_of[v] < 0) anchor_of[v] = a;
        if (v > 0) ++inv.nnpinv;
    }
    inv.node_of_col.assign(nnpgl, 0);
    for (int c = 0; c < nnpgl; ++c) {
        const int v = mygrad[c];
        if (v <= 0 || v > na35 || anchor_of[(v - 1) / nvinv + 1] < 0)
            FAIL(F25_ERR_ARG, "cannot find the anchor of a column");
        inv.node_of_col[c] = anchor_of[(v - 1) / nvinv + 1];
        const int len = jcsc[c + 1] - jcsc[c];
        if (len < 0) FAIL(F25_ERR_ARG, "JCSC_FDIST not ascending");
        inv.mbufrhs = std::max(inv.mbufrhs, len);
    }
    for (int i = 0; i < nz_fdist; ++i)
        if (icsc[i] < 1 || icsc[i] > ndof) FAIL(F25_ERR_ARG, "ICSC_FDIST out of range");
    for (size_t i = 0; i < inv.mconn.size(); ++i)
        if (inv.mconn[i] < 0 || inv.mconn[i] > nelem) FAIL(F25_ERR_ARG, "MCONN out of range");
    HIPCHK(hipSetDevice(m->dev));
    m->inv = std::move(inv);
    return inv_upload(m);
}

extern "C" int f25_inv_get(f25_mesh *m, int *maskg, int *mygrad, int *jcsc, int *icsc, double *ew) {
    if (!m || !m->have_inv) FAIL(F25_ERR_STATE, "f25_inv_setup has not been called");
    if (maskg) std::copy(m->inv.maskg.begin(), m->inv.maskg.end(), maskg);
    if (mygrad) std::copy(m->inv.mygrad.begin(), m->inv.mygrad.end(), mygrad);
    if (jcsc) std::copy(m->inv.jcsc.begin(), m->inv.jcsc.end(), jcsc);
    if (icsc) std::copy(m->inv.icsc.begin(), m->inv.icsc.end(), icsc);
    if (ew) std::copy(m->inv.elem_wts.begin(), m->inv.elem_wts.end(), ew);
    return 0;
}

// the element each column's product comes from: the last of MCONN(a,:)
// (VFORCE25 and QPMFREE clear the column per element, vforce25.f:252-256,
// qpmfree.f90:810-813, 1155-1160), and the anchor's corner in it (-1 when
// the element does not hold the anchor: the collocated field is then zero)
static int last_elements(const f25_mesh *m, std::vector<int> &ce, std::vector<int> &ck) {
    const InvInfo &inv = m->inv;
    const int n = inv.nnpgl, nnpg = m->d.nnpg;
    ce.assign(n, 0);
    ck.assign(n, -1);
    for (int c = 0; c < n; ++c) {
        const int a = inv.node_of_col[c];
        int e = 0;
        for (int k = 0; k < inv.ncon; ++k) {
            int v = inv.mconn[(size_t)k * nnpg + a];
            if (v <= 0) break;
            e = v;
        }
        if (e <= 0) FAIL(F25_ERR_STATE, "inversion anchor without elements");
        --e;
        ce[c] = e;
        for (int k = 0; k < 4; ++k)
            if (m->ieng[4 * e + k] - 1 == a) ck[c] = k;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Adjoint-source pipeline (handle-free: the receiver tables are the
// caller's).  Run on the current device.
static int check_mrdof(int nrec, int ndof, const int *mrdof, bool allow_out) {
    if (!mrdof) FAIL(F25_ERR_ARG, "null MRDOF");
    for (int k = 0; k < 3 * nrec; ++k)
        if (mrdof[k] < 0 || (!allow_out && mrdof[k] > ndof)) FAIL(F25_ERR_ARG, "MRDOF out of range");
    return 0;
}

static int exresp_check(int ndof, int nsrc, int nrec, const double *py, const double *yrec, const int *mrdof,
                        const void *recv) {
    if (ndof <= 0 || nsrc <= 0 || nrec <= 0 || !py || !yrec || !recv) FAIL(F25_ERR_ARG, "invalid arguments");
    return check_mrdof(nrec, ndof, mrdof, false);
}

static int exresp_core(int ndof, int nsrc, int nrec, double freq, const double *py, const double *yrec,
                       const float2 *sol_dev, const int *mrdof, const void *recv, float2 *est_dev, hipStream_t s) {
    double *dpy = nullptr, *dy = nullptr;
    int *dm = nullptr;
    float2 *dr = nullptr;
    int r = upload(&dpy, py, (size_t)nsrc, s);
    if (!r) r = upload(&dy, yrec, (size_t)nrec, s);
    if (!r) r = upload(&dm, mrdof, 3 * (size_t)nrec, s);
    if (!r) r = upload(&dr, (const float2 *)recv, 3 * (size_t)nrec, s);
    if (!r) {
        const int n = nsrc * nrec;
        hipLaunchKernelGGL(k_exresp, dim3((n + 255) / 256), dim3(256), 0, s, ndof, nsrc, nrec,
                           6.2831853071795862 * freq, dpy, dy, sol_dev, dm, dr, est_dev);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) { g_err = hipGetErrorString(e); r = F25_ERR_HIP; }
    }
    dfree(dpy); dfree(dy); dfree(dm); dfree(dr);
    return r;
}

extern "C" int f25_exresp25d_dev(int ndof, int nsrc, int nrec, double freq, const double *py, const double *yrec,
                                 const void *sol_dev, const int *mrdof, const void *recv, void *est_dev) {
    if (!sol_dev || !est_dev) FAIL(F25_ERR_ARG, "null argument");
    TRY(exresp_check(ndof, nsrc, nrec, py, yrec, mrdof, recv));
    TmpStream ts;
    HIPCHK(hipStreamCreate(&ts.s));
    return exresp_core(ndof, nsrc, nrec, freq, py, yrec, (const float2 *)sol_dev, mrdof, recv, (float2 *)est_dev,
                       ts.s);
}

extern "C" int f25_exresp25d(int ndof, int nsrc, int nrec, double freq, const double *py, const double *yrec,
                             const void *sol, const int *mrdof, const void *recv, void *est) {
    if (!sol || !est) FAIL(F25_ERR_ARG, "null argument");
    TRY(exresp_check(ndof, nsrc, nrec, py, yrec, mrdof, recv));
    TmpStream ts;
    HIPCHK(hipStreamCreate(&ts.s));
    float2 *ds = nullptr, *de = nullptr;
    const size_t ne = 3 * (size_t)nrec * nsrc;
    int r = upload(&ds, (const float2 *)sol, (size_t)ndof * nsrc, ts.s);
    if (!r) r = dalloc(&de, ne);
    if (!r) r = exresp_core(ndof, nsrc, nrec, freq, py, yrec, ds, mrdof, recv, de, ts.s);
    if (!r) {
        hipError_t e = hipMemcpy(est, de, ne * sizeof(float2), hipMemcpyDeviceToHost);
        if (e != hipSuccess) { g_err = hipGetErrorString(e); r = F25_ERR_HIP; }
    }
    dfree(ds); dfree(de);
    return r;
}

static int cadjlam_check(int ndof, int nfreq, int nrec, int nsrc, int npair, const int *ifreq, const int *isrc,
                         const double *freq, const double *py, const double *yrec, const int *mrdof, int *iwarn) {
    if (ndof <= 0 || nfreq <= 0 || nrec <= 0 || nsrc <= 0 || npair <= 0 || npair > 65535 || !ifreq || !isrc ||
        !freq || !py || !yrec || !iwarn)
        FAIL(F25_ERR_ARG, "invalid arguments");
    TRY(check_mrdof(nrec, ndof, mrdof, true));
    for (int p = 0; p < npair; ++p)
        if (ifreq[p] < 1 || ifreq[p] > nfreq || isrc[p] < 1 || isrc[p] > nsrc)
            FAIL(F25_ERR_ARG, "IFREQ or ISRC out of range");
    return 0;
}

// the data covariance (LCOVD): covd set, with ldwght and the nrec - 1
// normalised receiver spacings dx (checked by the caller)
struct CovdArgs {
    int covd, ldwght;
    const double *dx;
};

static int cadjlam_core(int ndof, int nfreq, int nrec, int nsrc, int npair, const int *ifreq, const int *isrc,
                        const double *freq, const double *py, double azmod, int norm, int irestp, float epss,
                        float dtmax, int ibphase, int lunwrap, const double *yrec, const int *mrdof,
                        const float2 *est_dev, const float2 *obs_dev, const float *w_dev, float2 *rhs_dev,
                        int *iwarn, hipStream_t s, CovdArgs cv = CovdArgs{0, 0, nullptr}) {
    AdjParams prm{(float)azmod, epss, dtmax, norm, irestp, ibphase, lunwrap ? 1 : 0};
    int *df = nullptr, *dsr = nullptr, *dm = nullptr, *dw = nullptr, *dinv = nullptr, *dpos = nullptr;
    double *dfr = nullptr, *dpy = nullptr, *dy = nullptr, *ddx = nullptr;
    float2 *dv = nullptr, *ddif = nullptr, *dres = nullptr;
    const size_t n3p = 3 * (size_t)nrec * npair;
    int r = upload(&df, ifreq, (size_t)npair, s);
    if (!r) r = upload(&dsr, isrc, (size_t)npair, s);
    if (!r) r = upload(&dfr, freq, (size_t)npair, s);
    if (!r) r = upload(&dpy, py, (size_t)npair, s);
    if (!r) r = upload(&dy, yrec, (size_t)nrec, s);
    if (!r) r = upload(&dm, mrdof, 3 * (size_t)nrec, s);
    if (!r) r = dalloc(&dv, n3p);
    if (!r) r = dalloc(&dw, 1);
    if (!r && cv.covd) {
        if (cv.ldwght && nrec > 1) r = upload(&ddx, cv.dx, (size_t)nrec - 1, s);
        else r = dalloc(&ddx, 1);
        if (!r) r = dalloc(&dinv, n3p);
        if (!r) r = dalloc(&dpos, n3p);
        if (!r) r = dalloc(&ddif, n3p);
        if (!r) r = dalloc(&dres, n3p);
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!r && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) {
        g_err = "hipEventCreate failed";
        r = F25_ERR_HIP;
    }
    if (!r) {
        fill_async(dw, (size_t)1, 0, s);
        (void)hipEventRecord(e0, s);
        fill_async(rhs_dev, (size_t)ndof * npair, make_float2(0.f, 0.f), s);
        if (cv.covd)
            hipLaunchKernelGGL(k_cadjlam_covd, dim3(npair), dim3(256), 0, s, ndof, nfreq, nrec, df, dsr, dfr, dpy,
                               prm, cv.ldwght, ddx, dy, dm, est_dev, obs_dev, w_dev, dinv, dpos, ddif, dres, dv,
                               rhs_dev, dw);
        else
            hipLaunchKernelGGL(k_cadjlam, dim3(npair), dim3(256), 0, s, ndof, nfreq, nrec, df, dsr, dfr, dpy, prm,
                               dy, dm, est_dev, obs_dev, w_dev, dv, rhs_dev, dw);
        (void)hipEventRecord(e1, s);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(iwarn, dw, sizeof(int), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) { g_err = hipGetErrorString(e); r = F25_ERR_HIP; }
        float ms = 0.f;
        if (!r && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) record_ms(cv.covd ? "cadjlam_covd" : "cadjlam", ms);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dfree(df); dfree(dsr); dfree(dfr); dfree(dpy); dfree(dy); dfree(dm); dfree(dv); dfree(dw);
    dfree(ddx); dfree(dinv); dfree(dpos); dfree(ddif); dfree(dres);
    return r;
}

extern "C" int f25_cadjlam_rhs_dev(int ndof, int nfreq, int nrec, int nsrc, int npair, const int *ifreq,
                                   const int *isrc, const double *freq, const double *py, double azmod, int norm,
                                   int irestp, float epss, float dtmax, int ibphase, int lunwrap,
                                   const double *yrec, const int *mrdof, const void *est_dev, const void *obs_dev,
                                   const float *wghts_dev, void *rhs_dev, int *iwarn) {
    if (!est_dev || !obs_dev || !wghts_dev || !rhs_dev) FAIL(F25_ERR_ARG, "null argument");
    TRY(cadjlam_check(ndof, nfreq, nrec, nsrc, npair, ifreq, isrc, freq, py, yrec, mrdof, iwarn));
    TmpStream ts;
    HIPCHK(hipStreamCreate(&ts.s));
    return cadjlam_core(ndof, nfreq, nrec, nsrc, npair, ifreq, isrc, freq, py, azmod, norm, irestp, epss, dtmax,
                        ibphase, lunwrap, yrec, mrdof, (const float2 *)est_dev, (const float2 *)obs_dev, wghts_dev,
                        (float2 *)rhs_dev, iwarn, ts.s);
}

extern "C" int f25_cadjlam_rhs(int ndof, int nfreq, int nrec, int nsrc, int npair, const int *ifreq,
                               const int *isrc, const double *freq, const double *py, double azmod, int norm,
                               int irestp, float epss, float dtmax, int ibphase, int lunwrap, const double *yrec,
                               const int *mrdof, const void *est, const void *obs, const float *wghts, void *rhs,
                               int *iwarn) {
    if (!est || !obs || !wghts || !rhs) FAIL(F25_ERR_ARG, "null argument");
    TRY(cadjlam_check(ndof, nfreq, nrec, nsrc, npair, ifreq, isrc, freq, py, yrec, mrdof, iwarn));
    TmpStream ts;
    HIPCHK(hipStreamCreate(&ts.s));
    const size_t nd = 3 * (size_t)nfreq * nrec * nsrc;
    float2 *de = nullptr, *dob = nullptr, *dr = nullptr;
    float *dw = nullptr;
    int r = upload(&de, (const float2 *)est, nd, ts.s);
    if (!r) r = upload(&dob, (const float2 *)obs, nd, ts.s);
    if (!r) r = upload(&dw, wghts, nd, ts.s);
    if (!r) r = dalloc(&dr, (size_t)ndof * npair);
    if (!r) r = cadjlam_core(ndof, nfreq, nrec, nsrc, npair, ifreq, isrc, freq, py, azmod, norm, irestp, epss,
                             dtmax, ibphase, lunwrap, yrec, mrdof, de, dob, dw, dr, iwarn, ts.s);
    if (!r) {
        hipError_t e = hipMemcpy(rhs, dr, (size_t)ndof * npair * sizeof(float2), hipMemcpyDeviceToHost);
        if (e != hipSuccess) { g_err = hipGetErrorString(e); r = F25_ERR_HIP; }
    }
    dfree(de); dfree(dob); dfree(dw); dfree(dr);
    return r;
}

// LAPLACE_1DV2 divides by the summed spacings: every spacing must be a
// positive finite number (REC_SPACE1D's are >= 1); the reference stops
// with an error when a zero sum is reached
static int covd_check(int nrec, int ldwght, const double *dx) {
    if (!ldwght || nrec < 2) return 0;
    if (!dx) FAIL(F25_ERR_ARG, "dx is null");
    for (int i = 0; i < nrec - 1; ++i)
        if (!(dx[i] > 0.0) || !std::isfinite(dx[i])) FAIL(F25_ERR_ARG, "receiver spacings must be positive and finite");
    return 0;
}

extern "C" int f25_cadjlam_rhs_covd_dev(int ndof, int nfreq, int nrec, int nsrc, int npair, const int *ifreq,
                                        const int *isrc, const double *freq, const double *py, double azmod,
                                        int irestp, float dtmax, int ibphase, int lunwrap, int ldwght,
                                        const double *dx, const double *yrec, const int *mrdof, const void *est_dev,
                                        const void *obs_dev, const float *wghts_dev, void *rhs_dev, int *iwarn) {
    if (!est_dev || !obs_dev || !wghts_dev || !rhs_dev) FAIL(F25_ERR_ARG, "null argument");
    TRY(cadjlam_check(ndof, nfreq, nrec, nsrc, npair, ifreq, isrc, freq, py, yrec, mrdof, iwarn));
    TRY(covd_check(nrec, ldwght, dx));
    TmpStream ts;
    HIPCHK(hipStreamCreate(&ts.s));
    return cadjlam_core(ndof, nfreq, nrec, nsrc, npair, ifreq, isrc, freq, py, azmod, 2, irestp, 0.f, dtmax,
                        ibphase, lunwrap, yrec, mrdof, (const float2 *)est_dev, (const float2 *)obs_dev, wghts_dev,
                        (float2 *)rhs_dev, iwarn, ts.s, CovdArgs{1, ldwght ? 1 : 0, dx});
}

extern "C" int f25_cadjlam_rhs_covd(int ndof, int nfreq, int nrec, int nsrc, int npair, const int *ifreq,
                                    const int *isrc, const double *freq, const double *py, double azmod, int irestp,
                                    float dtmax, int ibphase, int lunwrap, int ldwght, const double *dx,
                                    const double *yrec, const int *mrdof, const void *est, const void *obs,
                                    const float *wghts, void *rhs, int *iwarn) {
    if (!est || !obs || !wghts || !rhs) FAIL(F25_ERR_ARG, "null argument");
    TRY(cadjlam_check(ndof, nfreq, nrec, nsrc, npair, ifreq, isrc, freq, py, yrec, mrdof, iwarn));
    TRY(covd_check(nrec, ldwght, dx));
    TmpStream ts;
    HIPCHK(hipStreamCreate(&ts.s));
    const size_t nd = 3 * (size_t)nfreq * nrec * nsrc;
    float2 *de = nullptr, *dob = nullptr, *dr = nullptr;
    float *dw = nullptr;
    int r = upload(&de, (const float2 *)est, nd, ts.s);
    if (!r) r = upload(&dob, (const float2 *)obs, nd, ts.s);
    if (!r) r = upload(&dw, wghts, nd, ts.s);
    if (!r) r = dalloc(&dr, (size_t)ndof * npair);
    if (!r) r = cadjlam_core(ndof, nfreq, nrec, nsrc, npair, ifreq, isrc, freq, py, azmod, 2, irestp, 0.f, dtmax,
                             ibphase, lunwrap, yrec, mrdof, de, dob, dw, dr, iwarn, ts.s,
                             CovdArgs{1, ldwght ? 1 : 0, dx});
    if (!r) {
        hipError_t e = hipMemcpy(rhs, dr, (size_t)ndof * npair * sizeof(float2), hipMemcpyDeviceToHost);
        if (e != hipSuccess) { g_err = hipGetErrorString(e); r = F25_ERR_HIP; }
    }
    dfree(de); dfree(dob); dfree(dw); dfree(dr);
    return r;
}

// REC_SPACE1D (locrec2d.f:157-223): receiver spacings along the line,
// divided by the smallest; host arithmetic
extern "C" int f25_rec_space1d(int nrec, const double *xrec, double *dx) {
    if (nrec < 1 || !xrec || (nrec > 1 && !dx)) FAIL(F25_ERR_ARG, "invalid arguments");
    for (int i = 0; i < nrec - 1; ++i) dx[i] = 0.0;
    if (nrec == 1) FAIL(F25_ERR_ARG, "only one receiver");
    const double xdir = xrec[1] - xrec[0];
    double dxmin = HUGE_VAL;
    for (int i = 0; i < nrec - 1; ++i) {
        dx[i] = std::fabs(xrec[i + 1] - xrec[i]);
        dxmin = std::fmin(dxmin, dx[i]);
        const bool dup = dx[i] < 1e-10;
        const bool turn = i > 0 && ((xrec[i] > xrec[i - 1] && xdir < 0.0) || (xrec[i] < xrec[i - 1] && xdir > 0.0));
        if (dup || turn) {
            for (int j = 0; j < nrec - 1; ++j) dx[j] = 0.0;
            FAIL(F25_ERR_ARG, dup ? "duplicate receiver locations" : "receiver x not in order");
        }
    }
    for (int i = 0; i < nrec - 1; ++i) dx[i] = dx[i] / dxmin;
    return 0;
}

// ---------------------------------------------------------------------------
// QPMFREE's element products (CADJ_NU_ALPHA, MET225_MOD) on DABMKISOB25
// matrices.  2 Vp (PP) or 2 Vs (SS) at each column's anchor: DJACAB/DJACAB2's
// collocation constant (djac.f:94-103, 240-250).
static void dab_constants(const f25_mesh *m, std::vector<double> &da) {
    const InvInfo &inv = m->inv;
    da.resize(inv.nnpgl);
    for (int c = 0; c < inv.nnpgl; ++c) {
        const int a = inv.node_of_col[c];
        da[c] = col_is_p(inv, c) ? 2.0 * std::sqrt((m->lam[a] + 2.0 * m->mu[a]) / m->rho[a])
                                 : 2.0 * std::sqrt(m->mu[a] / m->rho[a]);
    }
}

// elements holding at least one local column on a corner (every element of
// MCONN(a,:) for every column's anchor a), coloured so that one launch per
// colour scatters without collisions
static int ensure_dab_part(f25_mesh *m) {
    InvInfo &inv = m->inv;
    if (!inv.dab_cptr.empty()) return 0;
    const int nelem = m->d.nelem, nnpg = m->d.nnpg;
    std::vector<int> col4((size_t)nelem * 4, -1);
    for (int c = 0; c < inv.nnpgl; ++c) {
        const int a = inv.node_of_col[c];
        if (inv.nvinv == 2) {   // PS: the P column stands for its pair (P, S)
            if ((inv.mygrad[c] - 1) % 2) continue;
            if (c + 1 >= inv.nnpgl || inv.node_of_col[c + 1] != a || inv.mygrad[c + 1] != inv.mygrad[c] + 1)
                FAIL(F25_ERR_ARG, "PS columns must come in (P, S) pairs (gengmask.f90:1078-1098)");
        }
        for (int k = 0; k < inv.ncon; ++k) {
            const int v = inv.mconn[(size_t)k * nnpg + a];
            if (v <= 0) break;
            const int e = v - 1;
            for (int kk = 0; kk < 4; ++kk)
                if (m->ieng[4 * e + kk] - 1 == a) {
                    if (col4[4 * e + kk] >= 0 && col4[4 * e + kk] != c)
                        FAIL(F25_ERR_ARG, "two local columns on one anchor");
                    col4[4 * e + kk] = c;
                }
        }
    }
    std::vector<int> list, sub;
    for (int e = 0; e < nelem; ++e)
        if (col4[4 * e] >= 0 || col4[4 * e + 1] >= 0 || col4[4 * e + 2] >= 0 || col4[4 * e + 3] >= 0) {
            list.push_back(e);
            sub.insert(sub.end(), m->ieng.begin() + 4 * e, m->ieng.begin() + 4 * e + 4);
        }
    std::vector<int> order, ptr;
    if (!list.empty()) colour_elements((int)list.size(), nnpg, sub.data(), order, ptr);
    else ptr.assign(1, 0);
    inv.dab_elem.resize(list.size());
    inv.dab_col4.resize(4 * list.size());
    for (size_t i = 0; i < list.size(); ++i) {
        const int e = list[order[i]];
        inv.dab_elem[i] = e;
        for (int k = 0; k < 4; ++k) inv.dab_col4[4 * i + k] = col4[4 * e + k];
    }
    inv.dab_cptr = ptr;
    return 0;
}

// LDS of k_dab_sf (independent of the number of vectors): the element's
// vectors (room for two), T1/T2 (whose space the projection's first-pass
// sums reuse), the projection fields A, B, C at the quadrature points and
// the 1-D tables v, v'  (65.4 KB at p6: two workgroups per CU)
static size_t dab_lds(const f25_mesh *m) {
    const int nee = 3 * m->tab.nen, nq = m->tab.nq, q1 = m->tab.nq1, n1 = m->tab.nl;
    return (size_t)2 * nee * sizeof(double2) + ((size_t)3 * q1 * n1 * 3 + 9 * (size_t)nq) * sizeof(double2) +
           (size_t)2 * n1 * q1 * sizeof(double);
}
// threads of a k_dab_sf workgroup: five waves at p6 (Q1 = 17)
static int dab_threads(const f25_mesh *m) { return m->tab.nq1 > 16 ? 320 : 256; }

// k_dab_sf for the mesh's order (the element loops are compiled per order)
typedef void (*DabKernel)(int, int, int, double, double, double, const int *, const double *, const double *,
                          const double *, const double *, const int *, const double *, const double *, const int *,
                          const float2 *, const float2 *, double2 *, double2 *, const float2 *, const float2 *,
                          const int *, const int *, const int *, float *, float *, float2 *, int *);
template <bool IMAG>
static DabKernel dab_kernel_t(int nl) {
    switch (nl) {
        case 3: return k_dab_sf<3, IMAG>;
        case 4: return k_dab_sf<4, IMAG>;
        case 5: return k_dab_sf<5, IMAG>;
        case 6: return k_dab_sf<6, IMAG>;
        case 7: return k_dab_sf<7, IMAG>;
        default: return nullptr;
    }
}
// DABMKISOB25 form (imaginary ky coupling) or DKM_VPVS form (real)
static DabKernel dab_kernel(int nl, bool imag = true) { return imag ? dab_kernel_t<true>(nl) : dab_kernel_t<false>(nl); }

// Positions of each column's element rows in its CSC pattern, for the
// column items of k_dab_sf (modes 1 and 2): the last element of the
// column's anchor, as the kernels take it.  A row absent from the pattern
// is the error the reference reports ('cant find inzrow', qpmfree.f90:819).
static int ensure_dab_pos(f25_mesh *m) {
    if (m->d_dab_pos) return 0;
    const InvInfo &inv = m->inv;
    const int nee = 3 * m->tab.nen;
    std::vector<int> ce, ck;
    TRY(last_elements(m, ce, ck));
    std::vector<int> pos((size_t)inv.nnpgl * nee, -1);
    for (int c = 0; c < inv.nnpgl; ++c) {
        const int *b = inv.icsc.data() + (inv.jcsc[c] - 1), *en = inv.icsc.data() + (inv.jcsc[c + 1] - 1);
        for (int j = 0; j < nee; ++j) {
            const int d = m->g.lm[(size_t)ce[c] * nee + j];
            if (d <= 0) continue;
            const int *f = std::lower_bound(b, en, d);
            if (f == en || *f != d) FAIL(F25_ERR_PATTERN, "an element row is missing from its column's CSC pattern");
            pos[(size_t)c * nee + j] = (int)(f - inv.icsc.data());
        }
    }
    TRY(upload(&m->d_dab_pos, pos.data(), pos.size(), m->st));
    HIPCHK(hipStreamSynchronize(m->st));
    return 0;
}

static int dab_prepare(f25_mesh *m) {
    if (!m->have_inv) FAIL(F25_ERR_STATE, "f25_inv_setup has not been called");
    if (!m->have_model) FAIL(F25_ERR_STATE, "f25_set_model has not been called");
    if (m->inv.nvinv != (m->inv.lisp && m->inv.liss ? 2 : 1)) FAIL(F25_ERR_ARG, "NVINV inconsistent with CINVTYPE");
    if (dab_lds(m) > 160 * 1024) FAIL(F25_ERR_ARG, "element order too high for the LDS tiling");
    HIPCHK(hipSetDevice(m->dev));
    if (!dab_kernel(m->tab.nl)) FAIL(F25_ERR_ARG, "unsupported order");
    if (dab_lds(m) > 64 * 1024)
        for (int im = 0; im < 2; ++im)
            HIPCHK(hipFuncSetAttribute((const void *)dab_kernel(m->tab.nl, im == 1),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)dab_lds(m)));
    return ensure_dab_pos(m);
}

// per-column items (mode 1): the last element, coefficient on its corner
static int dab_column_items(const f25_mesh *m, const std::vector<double> &da, const float *v,
                            std::vector<int> &elem, std::vector<double> &coef) {
    const InvInfo &inv = m->inv;
    std::vector<int> ce, ck;
    TRY(last_elements(m, ce, ck));
    elem = ce;
    coef.assign(8 * (size_t)inv.nnpgl, 0.0);
    for (int c = 0; c < inv.nnpgl; ++c) {
        if (ck[c] < 0) continue;
        float w = (float)inv.elem_wts[ce[c]];
        if (v) w *= v[inv.mygrad[c] - 1];   // CMPLX(SNGL(ELEM_WTS) * V(IA35)), qpmfree.f90:866
        const double cv = (double)w * da[c];
        // one field (PP, SS: the same coefficients for both); PS: the column's own
        if (inv.nvinv == 1 || col_is_p(inv, c)) coef[8 * c + ck[c]] = cv;
        if (inv.nvinv == 1 || !col_is_p(inv, c)) coef[8 * c + 4 + ck[c]] = cv;
    }
    return 0;
}

// kp, ks of k_dab_sf for DABMKISOB25's forms: P alone, S alone, or both
static void dab_kpks(const InvInfo &inv, double &kp, double &ks) {
    kp = inv.lisp ? 1.0 : 0.0;
    ks = inv.liss ? 1.0 : 0.0;
}

static int dab_launch_columns(f25_mesh *m, double kv, int nvec, const std::vector<int> &elem,
                              const std::vector<double> &coef, const float2 *x0, const float2 *x1,
                              const float2 *z0, const float2 *z1, float *p0, float *p1) {
    const InvInfo &inv = m->inv;
    int *de = nullptr;
    double *dc = nullptr;
    int r = upload(&de, elem.data(), elem.size(), m->st);
    if (!r) r = upload(&dc, coef.data(), coef.size(), m->st);
    if (!r) {
        double kp, ks;
        dab_kpks(inv, kp, ks);
        hipLaunchKernelGGL(dab_kernel(m->tab.nl), dim3(inv.nnpgl), dim3(dab_threads(m)), dab_lds(m), m->st, 1,
                           inv.nnpgl, nvec, kp, ks, kv, de, dc, m->d_t1d, m->d_wq, m->d_xq, m->d_ieng,
                           m->d_xl, m->d_zl, m->d_lm, x0, x1, (double2 *)nullptr, (double2 *)nullptr, z0, z1,
                           m->d_mygrad, m->d_jcsc, m->d_dab_pos, p0, p1, (float2 *)nullptr, m->d_err);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) { g_err = hipGetErrorString(e); r = F25_ERR_HIP; }
    }
    dfree(de); dfree(dc);
    return r;
}

static int dab_download(f25_mesh *m, float *dst, const float *src, size_t n) {
    hipError_t e = hipMemcpyAsync(dst, src, n * sizeof(float), hipMemcpyDeviceToHost, m->st);
    if (e != hipSuccess) FAIL(F25_ERR_HIP, hipGetErrorString(e));
    return 0;
}

// CFMAT_DIST (hessloc.f90:1-133 driving vforce25.f:1-369): per local
// column, the last connected element's DKM_VPVS product on WAVE, weighted by
// ELEM_WTS and written in CSC order.  DKM_VPVS's moduli are those of the
// derivative field DVP, DVS = DVP / VPVS^2 (PP) or DVP = VPVS^2 DVS (SS)
// (vforce25.f:164-179), DVP or DVS = 2 Vp or 2 Vs N_k (DJACAB).
extern "C" int f25_cfmat_dist(f25_mesh *m, double freq, double py, const void *wave, void *fmat) {
    if (!m || !wave || !fmat) FAIL(F25_ERR_ARG, "null argument");
    if (m->have_inv && m->inv.nvinv == 2)
        FAIL(F25_ERR_INVTYPE, "VFORCE25 has no joint P-S derivatives (vforce25.f:81-85)");
    TRY(dab_prepare(m));
    const InvInfo &inv = m->inv;
    const int n = inv.nnpgl;
    std::vector<int> ce, ck;
    TRY(last_elements(m, ce, ck));
    std::vector<double> coef(8 * (size_t)n, 0.0);
    const double v2 = inv.vpvs * inv.vpvs;
    for (int c = 0; c < n; ++c) {
        if (ck[c] < 0) continue;   // the element does not hold the anchor: zero column
        const int a = inv.node_of_col[c];
        const double vel = inv.lisp ? std::sqrt((m->lam[a] + 2.0 * m->mu[a]) / m->rho[a])
                                    : std::sqrt(m->mu[a] / m->rho[a]);
        coef[8 * c + ck[c]] = coef[8 * c + 4 + ck[c]] = inv.elem_wts[ce[c]] * 2.0 * vel;
    }
    const double kp = inv.lisp ? 1.0 : v2, ks = inv.lisp ? 1.0 / v2 : 1.0;
    const double kv = 6.2831853071795862 * freq * py;
    float2 *dw = nullptr, *df = nullptr;
    int *de = nullptr;
    double *dc = nullptr;
    int r = upload(&dw, (const float2 *)wave, m->g.ndof, m->st);
    if (!r) r = dalloc(&df, inv.nz_fdist);
    if (!r) r = upload(&de, ce.data(), ce.size(), m->st);
    if (!r) r = upload(&dc, coef.data(), coef.size(), m->st);
    if (!r) {
        (void)hipEventRecord(m->ev0, m->st);
        hipLaunchKernelGGL(dab_kernel(m->tab.nl, false), dim3(n), dim3(dab_threads(m)), dab_lds(m), m->st, 2, n, 1,
                           kp, ks, kv,
                           de, dc, m->d_t1d, m->d_wq, m->d_xq, m->d_ieng, m->d_xl, m->d_zl, m->d_lm, dw,
                           (const float2 *)nullptr, (double2 *)nullptr, (double2 *)nullptr, (const float2 *)nullptr,
                           (const float2 *)nullptr, m->d_mygrad, m->d_jcsc, m->d_dab_pos, (float *)nullptr,
                           (float *)nullptr, df, m->d_err);
        (void)hipEventRecord(m->ev1, m->st);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) { g_err = hipGetErrorString(e); r = F25_ERR_HIP; }
    }
    if (!r) r = check_err(m);
    float cms = 0.f;
    if (!r && hipEventElapsedTime(&cms, m->ev0, m->ev1) == hipSuccess) record_ms("cfmat", cms);
    if (!r) {
        hipError_t e = hipMemcpyAsync(fmat, df, inv.nz_fdist * sizeof(float2), hipMemcpyDeviceToHost, m->st);
        if (e == hipSuccess) e = hipStreamSynchronize(m->st);
        if (e != hipSuccess) { g_err = hipGetErrorString(e); r = F25_ERR_HIP; }
    }
    dfree(dw); dfree(df); dfree(de); dfree(dc);
    return r;
}

extern "C" int f25_qp_nu_alpha(f25_mesh *m, double freq, double py, const float *v, const void *adjlam,
                               const void *wave, void *rhsbnu, void *rhsalp, float *p3) {
    if (!m || !v || !adjlam || !wave || !rhsbnu || !rhsalp || !p3) FAIL(F25_ERR_ARG, "null argument");
    TRY(dab_prepare(m));
    TRY(ensure_dab_part(m));
    const InvInfo &inv = m->inv;
    const int ndof = m->g.ndof;
    const double kv = 6.2831853071795862 * freq * py;
    std::vector<double> da;
    dab_constants(m, da);
    // element items: coefficient of every corner that holds a local column
    // (PS: the corner's P column c and S column c + 1, each with its own V)
    const size_t ne = inv.dab_elem.size();
    std::vector<double> ecoef(8 * std::max<size_t>(ne, 1), 0.0);
    for (size_t i = 0; i < ne; ++i)
        for (int k = 0; k < 4; ++k) {
            const int c = inv.dab_col4[4 * i + k];
            if (c < 0) continue;
            const float ew = (float)inv.elem_wts[inv.dab_elem[i]];
            const double cp = (double)(ew * v[inv.mygrad[c] - 1]) * da[c];
            if (inv.nvinv == 2) {
                ecoef[8 * i + k] = cp;
                ecoef[8 * i + 4 + k] = (double)(ew * v[inv.mygrad[c + 1] - 1]) * da[c + 1];
            } else {
                ecoef[8 * i + k] = ecoef[8 * i + 4 + k] = cp;
            }
        }
    std::vector<int> celem;
    std::vector<double> ccoef;
    TRY(dab_column_items(m, da, v, celem, ccoef));
    float2 *du = nullptr, *dl = nullptr;
    double2 *ya = nullptr, *yb = nullptr;
    int *de = nullptr;
    double *dc = nullptr;
    float *dp3 = nullptr;
    int r = upload(&du, (const float2 *)wave, ndof, m->st);
    if (!r) r = upload(&dl, (const float2 *)adjlam, ndof, m->st);
    if (!r) r = dalloc(&ya, ndof);
    if (!r) r = dalloc(&yb, ndof);
    if (!r) r = upload(&de, inv.dab_elem.data(), ne, m->st);
    if (!r) r = upload(&dc, ecoef.data(), 8 * ne, m->st);
    if (!r) r = dalloc(&dp3, inv.na35);
    if (!r) {
        fill_async(ya, (size_t)ndof, make_double2(0.0, 0.0), m->st);
        fill_async(yb, (size_t)ndof, make_double2(0.0, 0.0), m->st);
        fill_async(dp3, (size_t)inv.na35, 0.f, m->st);
        (void)hipEventRecord(m->ev0, m->st);
        double kp, ks;
        dab_kpks(inv, kp, ks);
        for (size_t col = 0; col + 1 < inv.dab_cptr.size(); ++col) {
            const int o = inv.dab_cptr[col], n = inv.dab_cptr[col + 1] - o;
            if (n <= 0) continue;
            hipLaunchKernelGGL(dab_kernel(m->tab.nl), dim3(n), dim3(dab_threads(m)), dab_lds(m), m->st, 0, n, 2,
                               kp, ks, kv, de + o, dc + 8 * (size_t)o,
                               m->d_t1d, m->d_wq, m->d_xq, m->d_ieng, m->d_xl, m->d_zl, m->d_lm, du, dl, ya, yb,
                               (const float2 *)nullptr, (const float2 *)nullptr, m->d_mygrad, m->d_jcsc,
                               m->d_dab_pos, (float *)nullptr, (float *)nullptr, (float2 *)nullptr, m->d_err);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) { g_err = hipGetErrorString(e); r = F25_ERR_HIP; }
        if (!r) r = dab_launch_columns(m, kv, 1, celem, ccoef, du, nullptr, dl, nullptr, dp3, nullptr);
        (void)hipEventRecord(m->ev1, m->st);
    }
    if (!r) r = check_err(m);
    float ms = 0.f;
    if (!r && hipEventElapsedTime(&ms, m->ev0, m->ev1) == hipSuccess) record_ms("qp", ms);
    if (!r) {
        std::vector<double2> ha(ndof), hb(ndof);
        hipError_t e = hipMemcpyAsync(ha.data(), ya, ndof * sizeof(double2), hipMemcpyDeviceToHost, m->st);
        if (e == hipSuccess) e = hipMemcpyAsync(hb.data(), yb, ndof * sizeof(double2), hipMemcpyDeviceToHost, m->st);
        if (e == hipSuccess) r = dab_download(m, p3, dp3, inv.na35);
        if (e == hipSuccess && !r) e = hipStreamSynchronize(m->st);
        if (e != hipSuccess) { g_err = hipGetErrorString(e); r = F25_ERR_HIP; }
        if (!r) {
            float2 *oa = (float2 *)rhsalp, *ob = (float2 *)rhsbnu;
            for (int i = 0; i < ndof; ++i) {
                oa[i] = make_float2((float)ha[i].x, (float)ha[i].y);
                ob[i] = make_float2((float)hb[i].x, (float)hb[i].y);
            }
        }
    }
    dfree(du); dfree(dl); dfree(ya); dfree(yb); dfree(de); dfree(dc); dfree(dp3);
    return r;
}

extern "C" int f25_met225(f25_mesh *m, double freq, double py, const void *adjlam, const void *nu,
                          const void *alpha, const void *wave, float *p1, float *p2) {
    if (!m || !adjlam || !nu || !alpha || !wave || !p1 || !p2) FAIL(F25_ERR_ARG, "null argument");
    TRY(dab_prepare(m));
    const InvInfo &inv = m->inv;
    const int ndof = m->g.ndof;
    const double kv = 6.2831853071795862 * freq * py;
    std::vector<double> da;
    dab_constants(m, da);
    std::vector<int> celem;
    std::vector<double> ccoef;
    TRY(dab_column_items(m, da, nullptr, celem, ccoef));
    float2 *dl = nullptr, *dn = nullptr, *da_ = nullptr, *du = nullptr;
    float *d1 = nullptr, *d2 = nullptr;
    int r = upload(&dl, (const float2 *)adjlam, ndof, m->st);
    if (!r) r = upload(&dn, (const float2 *)nu, ndof, m->st);
    if (!r) r = upload(&da_, (const float2 *)alpha, ndof, m->st);
    if (!r) r = upload(&du, (const float2 *)wave, ndof, m->st);
    if (!r) r = dalloc(&d1, inv.na35);
    if (!r) r = dalloc(&d2, inv.na35);
    if (!r) {
        fill_async(d1, (size_t)inv.na35, 0.f, m->st);
        fill_async(d2, (size_t)inv.na35, 0.f, m->st);
        (void)hipEventRecord(m->ev0, m->st);
        r = dab_launch_columns(m, kv, 2, celem, ccoef, du, da_, dn, dl, d1, d2);
        (void)hipEventRecord(m->ev1, m->st);
    }
    if (!r) r = check_err(m);
    float ms = 0.f;
    if (!r && hipEventElapsedTime(&ms, m->ev0, m->ev1) == hipSuccess) record_ms("met225", ms);
    if (!r) r = dab_download(m, p1, d1, inv.na35);
    if (!r) r = dab_download(m, p2, d2, inv.na35);
    if (!r) {
        hipError_t e = hipStreamSynchronize(m->st);
        if (e != hipSuccess) { g_err = hipGetErrorString(e); r = F25_ERR_HIP; }
    }
    dfree(dl); dfree(dn); dfree(da_); dfree(du); dfree(d1); dfree(d2);
    return r;
}

extern "C" int f25_adjflam_dist(f25_mesh *m, const void *fmat, const void *adjlam, float *gprod) {
    if (!m || !fmat || !adjlam || !gprod) FAIL(F25_ERR_ARG, "null argument");
    if (!m->have_inv) FAIL(F25_ERR_STATE, "f25_inv_setup has not been called");
    HIPCHK(hipSetDevice(m->dev));
    float2 *df = nullptr, *dl = nullptr;
    float *dg = nullptr;
    TRY(upload(&df, (const float2 *)fmat, m->inv.nz_fdist, m->st));
    TRY(upload(&dl, (const float2 *)adjlam, m->g.ndof, m->st));
    TRY(dalloc(&dg, m->inv.na35));
    fill_async(dg, (size_t)m->inv.na35, 0.f, m->st);
    hipLaunchKernelGGL(k_adjflam, dim3((m->inv.nnpgl + 3) / 4), dim3(256), 0, m->st, m->inv.nnpgl,
                       m->d_mygrad, m->d_jcsc, m->d_icsc, df, dl, dg);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(gprod, dg, m->inv.na35 * sizeof(float), hipMemcpyDeviceToHost, m->st);
    if (e == hipSuccess) e = hipStreamSynchronize(m->st);
    dfree(df); dfree(dl); dfree(dg);
    if (e != hipSuccess) FAIL(F25_ERR_HIP, hipGetErrorString(e));
    return 0;
}

// ---------------------------------------------------------------------------
// Handle-free forms for callers that hold only the sparse structures
// (ASMB_PEFF and ADJFLAM_DIST receive no mesh).  They run on the current
// device, on a stream of their own, and allocate per call.
extern "C" int f25_asmb_peff_csr(int ndof, int ndofl, int nzloc, const char *cnp, const int *mydofs,
                                 const int *irptr_loc, const int *jcptr_loc, const void *cdist,
                                 const void *ue, void *buff) {
    if (ndof <= 0 || ndofl < 0 || ndofl > ndof || nzloc < 0 || !cnp || !mydofs || !irptr_loc ||
        !jcptr_loc || !cdist || !ue || !buff)
        FAIL(F25_ERR_ARG, "invalid arguments");
    if (irptr_loc[0] != 1 || irptr_loc[ndofl] - 1 != nzloc) FAIL(F25_ERR_ARG, "IRPTR_LOC inconsistent with NZLOC");
    std::vector<int64_t> ir(ndofl + 1);
    for (int i = 0; i <= ndofl; ++i) {
        ir[i] = (int64_t)irptr_loc[i] - 1;
        if (i && ir[i] < ir[i - 1]) FAIL(F25_ERR_ARG, "IRPTR_LOC not ascending");
    }
    std::vector<int> jc(nzloc);
    for (int i = 0; i < nzloc; ++i) {
        if (jcptr_loc[i] < 1 || jcptr_loc[i] > ndof) FAIL(F25_ERR_ARG, "JCPTR_LOC out of range");
        jc[i] = jcptr_loc[i] - 1;
    }
    for (int i = 0; i < ndofl; ++i)
        if (mydofs[i] < 1 || mydofs[i] > ndof) FAIL(F25_ERR_ARG, "MYDOFS out of range");
    TmpStream ts;
    HIPCHK(hipStreamCreate(&ts.s));
    int64_t *d_ir = nullptr;
    int *d_jc = nullptr, *d_my = nullptr;
    char *d_cnp = nullptr;
    float2 *ds = nullptr, *du = nullptr, *dp = nullptr;
    int r = 0;
    r = upload(&d_ir, ir.data(), ir.size(), ts.s);
    if (!r) r = upload(&d_jc, jc.data(), jc.size(), ts.s);
    if (!r) r = upload(&d_my, mydofs, (size_t)ndofl, ts.s);
    if (!r) r = upload(&d_cnp, cnp, (size_t)ndof, ts.s);
    if (!r) r = upload(&ds, (const float2 *)cdist, (size_t)nzloc, ts.s);
    if (!r) r = upload(&du, (const float2 *)ue, (size_t)ndof, ts.s);
    if (!r) r = dalloc(&dp, (size_t)ndof);
    if (!r) {
        fill_async(dp, (size_t)ndof, make_float2(0.f, 0.f), ts.s);
        if (ndofl > 0)
            hipLaunchKernelGGL(k_peff, dim3((ndofl + 3) / 4), dim3(256), 0, ts.s, ndofl, d_my, d_ir, d_jc,
                               d_cnp, ds, du, dp);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(buff, dp, ndof * sizeof(float2), hipMemcpyDeviceToHost, ts.s);
        if (e == hipSuccess) e = hipStreamSynchronize(ts.s);
        if (e != hipSuccess) { g_err = hipGetErrorString(e); r = F25_ERR_HIP; }
    }
    dfree(d_ir); dfree(d_jc); dfree(d_my); dfree(d_cnp); dfree(ds); dfree(du); dfree(dp);
    return r;
}

extern "C" int f25_adjflam_csc(int nnpgl, int ndof, int nz_fdist, int na35, const int *mygrad,
                               const int *jcsc, const int *icsc, const void *fmat, const void *adjlam,
                               float *vwork) {
    if (nnpgl < 0 || ndof <= 0 || nz_fdist < 0 || na35 <= 0 || !mygrad || !jcsc || !icsc || !fmat ||
        !adjlam || !vwork)
        FAIL(F25_ERR_ARG, "invalid arguments");
    if (jcsc[0] != 1 || jcsc[nnpgl] - 1 != nz_fdist) FAIL(F25_ERR_ARG, "JCSC_FDIST inconsistent with NZ_FDIST");
    std::vector<char> seen(na35, 0);
    for (int c = 0; c < nnpgl; ++c) {
        if (mygrad[c] < 1 || mygrad[c] > na35) FAIL(F25_ERR_ARG, "MYGRAD out of range");
        if (seen[mygrad[c] - 1]) FAIL(F25_ERR_ARG, "MYGRAD repeats a gradient point");
        seen[mygrad[c] - 1] = 1;
        if (jcsc[c + 1] < jcsc[c]) FAIL(F25_ERR_ARG, "JCSC_FDIST not ascending");
    }
    for (int i = 0; i < nz_fdist; ++i)
        if (icsc[i] < 1 || icsc[i] > ndof) FAIL(F25_ERR_ARG, "ICSC_FDIST out of range");
    TmpStream ts;
    HIPCHK(hipStreamCreate(&ts.s));
    int *d_my = nullptr, *d_jc = nullptr, *d_ic = nullptr;
    float2 *df = nullptr, *dl = nullptr;
    float *dg = nullptr;
    int r = upload(&d_my, mygrad, (size_t)nnpgl, ts.s);
    if (!r) r = upload(&d_jc, jcsc, (size_t)nnpgl + 1, ts.s);
    if (!r) r = upload(&d_ic, icsc, (size_t)nz_fdist, ts.s);
    if (!r) r = upload(&df, (const float2 *)fmat, (size_t)nz_fdist, ts.s);
    if (!r) r = upload(&dl, (const float2 *)adjlam, (size_t)ndof, ts.s);
    if (!r) r = dalloc(&dg, (size_t)na35);
    if (!r) {
        fill_async(dg, (size_t)na35, 0.f, ts.s);
        if (nnpgl > 0)
            hipLaunchKernelGGL(k_adjflam, dim3((nnpgl + 3) / 4), dim3(256), 0, ts.s, nnpgl, d_my, d_jc, d_ic,
                               df, dl, dg);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(vwork, dg, na35 * sizeof(float), hipMemcpyDeviceToHost, ts.s);
        if (e == hipSuccess) e = hipStreamSynchronize(ts.s);
        if (e != hipSuccess) { g_err = hipGetErrorString(e); r = F25_ERR_HIP; }
    }
    dfree(d_my); dfree(d_jc); dfree(d_ic); dfree(df); dfree(dl); dfree(dg);
    return r;
}

// ---------------------------------------------------------------------------
// Jacobians (PCHADJ) and the Gauss-Newton contraction (GNMFREE)
static int pchadj_core(f25_mesh *m, int nrhs, const float2 *gt, const float2 *f, float2 *jac, float *hd) {
    const int n = m->inv.nnpgl;
    HIPCHK(hipEventRecord(m->ev0, m->st));
    hipLaunchKernelGGL(k_pchadj, dim3((n + 3) / 4), dim3(256), 0, m->st, n, nrhs, m->d_mygrad, m->d_jcsc,
                       m->d_icsc, gt, f, jac, hd);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(m->ev1, m->st));
    HIPCHK(hipStreamSynchronize(m->st));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, m->ev0, m->ev1));
    record_ms("pchadj", ms);
    return 0;
}

extern "C" int f25_pchadj_jac_dev(f25_mesh *m, int nrhs, const void *gt, const void *fmat, void *jac,
                                  float *hdiag) {
    if (!m || nrhs <= 0 || !gt || !fmat || !jac || !hdiag) FAIL(F25_ERR_ARG, "invalid arguments");
    if (!m->have_inv) FAIL(F25_ERR_STATE, "inversion structures missing (f25_inv_setup / f25_inv_attach)");
    HIPCHK(hipSetDevice(m->dev));
    return pchadj_core(m, nrhs, (const float2 *)gt, (const float2 *)fmat, (float2 *)jac, hdiag);
}

extern "C" int f25_pchadj_jac(f25_mesh *m, int nrhs, int lrhs, const void *grnsfns, const void *fmat,
                              void *jac, float *hdiag) {
    if (!m || nrhs <= 0 || !grnsfns || !fmat || !jac || !hdiag) FAIL(F25_ERR_ARG, "invalid arguments");
    if (!m->have_inv) FAIL(F25_ERR_STATE, "inversion structures missing (f25_inv_setup / f25_inv_attach)");
    if (lrhs < m->g.ndof) FAIL(F25_ERR_ARG, "LRHS smaller than NDOF");
    HIPCHK(hipSetDevice(m->dev));
    const InvInfo &inv = m->inv;
    float2 *dg = nullptr, *dgt = nullptr, *df = nullptr, *dj = nullptr;
    float *dh = nullptr;
    const size_t ng = (size_t)nrhs * lrhs;
    int r = upload(&dg, (const float2 *)grnsfns, ng, m->st);
    if (!r) r = dalloc(&dgt, ng);
    if (!r) r = transpose_dev(m->st, (const float2 *)dg, dgt, nrhs, lrhs);
    if (!r) r = upload(&df, (const float2 *)fmat, (size_t)inv.nz_fdist, m->st);
    if (!r) r = dalloc(&dj, (size_t)nrhs * inv.nnpgl);
    if (!r) r = upload(&dh, hdiag, (size_t)inv.na35, m->st);
    if (!r) r = pchadj_core(m, nrhs, dgt, df, dj, dh);
    if (!r) {
        hipError_t e = hipMemcpyAsync(jac, dj, (size_t)nrhs * inv.nnpgl * sizeof(float2), hipMemcpyDeviceToHost, m->st);
        if (e == hipSuccess) e = hipMemcpyAsync(hdiag, dh, (size_t)inv.na35 * sizeof(float), hipMemcpyDeviceToHost, m->st);
        if (e == hipSuccess) e = hipStreamSynchronize(m->st);
        if (e != hipSuccess) { g_err = hipGetErrorString(e); r = F25_ERR_HIP; }
    }
    dfree(dg); dfree(dgt); dfree(df); dfree(dj); dfree(dh);
    return r;
}

// packed offsets of the njac Jacobians (nobs[j] x nnpgl each) and of their
// JP vectors
static int gn_offsets(int njac, const int *nobs, int nnpgl, std::vector<int64_t> &joff,
                      std::vector<int64_t> &poff, int64_t &jt, int64_t &pt) {
    joff.resize(njac);
    poff.resize(njac);
    jt = pt = 0;
    for (int j = 0; j < njac; ++j) {
        if (nobs[j] < 0) FAIL(F25_ERR_ARG, "negative observation count");
        joff[j] = jt;
        poff[j] = pt;
        jt += (int64_t)nobs[j] * nnpgl;
        pt += nobs[j];
    }
    return 0;
}

static int gn_check(int nnpgl, int na35, const int *mygrad, int njac) {
    if (nnpgl <= 0 || na35 <= 0 || !mygrad || njac <= 0 || njac > 65535) FAIL(F25_ERR_ARG, "invalid arguments");
    std::vector<char> seen(na35, 0);
    for (int c = 0; c < nnpgl; ++c) {
        if (mygrad[c] < 1 || mygrad[c] > na35) FAIL(F25_ERR_ARG, "MYGRAD out of range");
        if (seen[mygrad[c] - 1]) FAIL(F25_ERR_ARG, "MYGRAD repeats a gradient point");
        seen[mygrad[c] - 1] = 1;
    }
    return 0;
}

// J p in two deterministic passes: column chunks of J_j into partial sums
// (at most 256 chunks: part holds nch * pt double2), then the chunks added
// in order.
static void gn_jp_chunks(int nnpgl, int &nch, int &cch) {
    nch = std::min(256, (nnpgl + 63) / 64);
    cch = (nnpgl + nch - 1) / nch;
    nch = (nnpgl + cch - 1) / cch;
}

static void launch_gn_jp(hipStream_t st, int nnpgl, int njac, int64_t pt, const int *dn, const int64_t *djo,
                         const int64_t *dpo, const int *dmy, const float2 *jac, const double *p, double2 *part,
                         double2 *jp) {
    int nch, cch;
    gn_jp_chunks(nnpgl, nch, cch);
    hipLaunchKernelGGL(k_gn_jp, dim3(nch, njac), dim3(256), 0, st, nnpgl, cch, pt, dn, djo, dpo, dmy, jac, p, part);
    const int nb = (int)std::min<int64_t>((pt + 255) / 256, 4096);
    hipLaunchKernelGGL(k_gn_jsum, dim3(std::max(nb, 1)), dim3(256), 0, st, nch, pt, part, jp);
}

static int gnmfree_core(f25_mesh *m, int njac, const int *nobs, const double *alpha, const float2 *jac,
                        const double *p, double *z) {
    const int n = m->inv.nnpgl;
    std::vector<int64_t> joff, poff;
    int64_t jt, pt;
    TRY(gn_offsets(njac, nobs, n, joff, poff, jt, pt));
    int *dn = nullptr;
    int64_t *djo = nullptr, *dpo = nullptr;
    double *da = nullptr;
    double2 *djp = nullptr, *dpart = nullptr;
    int nch, cch;
    gn_jp_chunks(n, nch, cch);
    int r = upload(&dn, nobs, (size_t)njac, m->st);
    if (!r) r = upload(&djo, joff.data(), (size_t)njac, m->st);
    if (!r) r = upload(&dpo, poff.data(), (size_t)njac, m->st);
    if (!r) r = upload(&da, alpha, (size_t)njac, m->st);
    if (!r) r = dalloc(&djp, (size_t)std::max<int64_t>(pt, 1));
    if (!r) r = dalloc(&dpart, (size_t)nch * std::max<int64_t>(pt, 1));
    if (!r) {
        (void)hipEventRecord(m->ev0, m->st);
        launch_gn_jp(m->st, n, njac, pt, dn, djo, dpo, m->d_mygrad, jac, p, dpart, djp);
        hipLaunchKernelGGL(k_gn_z, dim3((n + 3) / 4), dim3(256), 0, m->st, n, njac, dn, djo, dpo, da,
                           m->d_mygrad, jac, djp, z);
        (void)hipEventRecord(m->ev1, m->st);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(m->st);
        if (e != hipSuccess) { g_err = hipGetErrorString(e); r = F25_ERR_HIP; }
        float ms = 0.f;
        if (!r && hipEventElapsedTime(&ms, m->ev0, m->ev1) == hipSuccess) record_ms("gnmfree", ms);
    }
    dfree(dn); dfree(djo); dfree(dpo); dfree(da); dfree(djp); dfree(dpart);
    return r;
}

// GNMFREE in two halves on caller-held arrays, for a caller that reduces JP
// across ranks in between (gnmfree.f90:195-196)
extern "C" int f25_gn_jp(int nnpgl, int na35, const int *mygrad, int njac, const int *nobs, const void *jac,
                         const double *p, void *jp) {
    TRY(gn_check(nnpgl, na35, mygrad, njac));
    if (!nobs || !jac || !p || !jp) FAIL(F25_ERR_ARG, "invalid arguments");
    std::vector<int64_t> joff, poff;
    int64_t jt, pt;
    TRY(gn_offsets(njac, nobs, nnpgl, joff, poff, jt, pt));
    TmpStream ts;
    HIPCHK(hipStreamCreate(&ts.s));
    int *dmy = nullptr, *dn = nullptr;
    int64_t *djo = nullptr, *dpo = nullptr;
    float2 *dj = nullptr;
    double *dp_ = nullptr;
    double2 *djp = nullptr, *dpart = nullptr;
    int nch, cch;
    gn_jp_chunks(nnpgl, nch, cch);
    int r = upload(&dmy, mygrad, (size_t)nnpgl, ts.s);
    if (!r) r = upload(&dn, nobs, (size_t)njac, ts.s);
    if (!r) r = upload(&djo, joff.data(), (size_t)njac, ts.s);
    if (!r) r = upload(&dpo, poff.data(), (size_t)njac, ts.s);
    if (!r) r = upload(&dj, (const float2 *)jac, (size_t)jt, ts.s);
    if (!r) r = upload(&dp_, p, (size_t)na35, ts.s);
    if (!r) r = dalloc(&djp, (size_t)std::max<int64_t>(pt, 1));
    if (!r) r = dalloc(&dpart, (size_t)nch * std::max<int64_t>(pt, 1));
    if (!r) {
        launch_gn_jp(ts.s, nnpgl, njac, pt, dn, djo, dpo, dmy, dj, dp_, dpart, djp);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess && pt > 0) e = hipMemcpyAsync(jp, djp, (size_t)pt * sizeof(double2), hipMemcpyDeviceToHost, ts.s);
        if (e == hipSuccess) e = hipStreamSynchronize(ts.s);
        if (e != hipSuccess) { g_err = hipGetErrorString(e); r = F25_ERR_HIP; }
    }
    dfree(dmy); dfree(dn); dfree(djo); dfree(dpo); dfree(dj); dfree(dp_); dfree(djp); dfree(dpart);
    return r;
}

extern "C" int f25_gn_jhz(int nnpgl, int na35, const int *mygrad, int njac, const int *nobs, const double *alpha,
                          const void *jac, const void *jp, double *z) {
    TRY(gn_check(nnpgl, na35, mygrad, njac));
    if (!nobs || !alpha || !jac || !jp || !z) FAIL(F25_ERR_ARG, "invalid arguments");
    std::vector<int64_t> joff, poff;
    int64_t jt, pt;
    TRY(gn_offsets(njac, nobs, nnpgl, joff, poff, jt, pt));
    TmpStream ts;
    HIPCHK(hipStreamCreate(&ts.s));
    int *dmy = nullptr, *dn = nullptr;
    int64_t *djo = nullptr, *dpo = nullptr;
    float2 *dj = nullptr;
    double *da = nullptr, *dz = nullptr;
    double2 *djp = nullptr;
    int r = upload(&dmy, mygrad, (size_t)nnpgl, ts.s);
    if (!r) r = upload(&dn, nobs, (size_t)njac, ts.s);
    if (!r) r = upload(&djo, joff.data(), (size_t)njac, ts.s);
    if (!r) r = upload(&dpo, poff.data(), (size_t)njac, ts.s);
    if (!r) r = upload(&da, alpha, (size_t)njac, ts.s);
    if (!r) r = upload(&dj, (const float2 *)jac, (size_t)jt, ts.s);
    if (!r) r = upload(&djp, (const double2 *)jp, (size_t)pt, ts.s);
    if (!r) r = upload(&dz, z, (size_t)na35, ts.s);
    if (!r) {
        hipLaunchKernelGGL(k_gn_z, dim3((nnpgl + 3) / 4), dim3(256), 0, ts.s, nnpgl, njac, dn, djo, dpo, da, dmy,
                           dj, djp, dz);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(z, dz, (size_t)na35 * sizeof(double), hipMemcpyDeviceToHost, ts.s);
        if (e == hipSuccess) e = hipStreamSynchronize(ts.s);
        if (e != hipSuccess) { g_err = hipGetErrorString(e); r = F25_ERR_HIP; }
    }
    dfree(dmy); dfree(dn); dfree(djo); dfree(dpo); dfree(da); dfree(dj); dfree(djp); dfree(dz);
    return r;
}

extern "C" int f25_gnmfree_dev(f25_mesh *m, int njac, const int *nobs, const double *alpha, const void *jac,
                               const double *p, double *z) {
    if (!m || njac <= 0 || njac > 65535 || !nobs || !alpha || !jac || !p || !z) FAIL(F25_ERR_ARG, "invalid arguments");
    if (!m->have_inv) FAIL(F25_ERR_STATE, "inversion structures missing (f25_inv_setup / f25_inv_attach)");
    HIPCHK(hipSetDevice(m->dev));
    return gnmfree_core(m, njac, nobs, alpha, (const float2 *)jac, p, z);
}

// ---- Gauss-Newton Hessian over the local columns ---------------------------
extern "C" int f25_gn_hessian_dev(f25_mesh *m, int njac, const int *nobs, const double *alpha, const void *jac,
                                  double *h) {
    if (!m || njac <= 0 || !nobs || !alpha || !jac || !h) FAIL(F25_ERR_ARG, "invalid arguments");
    if (!m->have_inv) FAIL(F25_ERR_STATE, "inversion structures missing (f25_inv_setup / f25_inv_attach)");
    for (int j = 0; j < njac; ++j)
        if (!(alpha[j] >= 0.0)) FAIL(F25_ERR_ARG, "alpha must be non-negative");
    HIPCHK(hipSetDevice(m->dev));
    const int n = m->inv.nnpgl;
    std::vector<int64_t> joff, poff;
    int64_t jt, pt;
    TRY(gn_offsets(njac, nobs, n, joff, poff, jt, pt));
    if (!m->blas) {
        if (rocblas_create_handle(&m->blas) != rocblas_status_success) {
            m->blas = nullptr;
            FAIL(F25_ERR_HIP, "rocblas_create_handle failed");
        }
    }
    if (rocblas_set_stream(m->blas, m->st) != rocblas_status_success) FAIL(F25_ERR_HIP, "rocblas_set_stream failed");
    // Jacobians in groups whose stacked rows fit a 1 GiB FP64 buffer; one
    // SYRK (H += A^T A, upper triangle) per group
    const int64_t cap_rows = std::max<int64_t>(2 * (int64_t)*std::max_element(nobs, nobs + njac),
                                               (int64_t(1) << 27) / std::max(n, 1));
    double *a = nullptr;
    TRY(dalloc(&a, (size_t)cap_rows * n));
    const float2 *J = static_cast<const float2 *>(jac);
    const double one = 1.0;
    int r = 0;
    (void)hipEventRecord(m->ev0, m->st);
    for (int j0 = 0; j0 < njac && !r;) {
        int64_t k = 0;
        int j1 = j0;
        while (j1 < njac && k + 2 * nobs[j1] <= cap_rows) k += 2 * nobs[j1++];
        int64_t roff = 0;
        for (int j = j0; j < j1; ++j) {
            if (!nobs[j]) continue;
            const int64_t cnt = (int64_t)nobs[j] * n;
            const int nb = (int)std::min<int64_t>((cnt + 255) / 256, 16384);
            hipLaunchKernelGGL(k_jac_rows, dim3(nb), dim3(256), 0, m->st, n, nobs[j], std::sqrt(alpha[j]),
                               J + joff[j], a, k, roff);
            roff += 2 * nobs[j];
        }
        if (k > 0 && rocblas_dsyrk(m->blas, rocblas_fill_upper, rocblas_operation_transpose, n, (rocblas_int)k, &one,
                                   a, (rocblas_int)k, &one, h, n) != rocblas_status_success) {
            g_err = "rocblas_dsyrk failed";
            r = F25_ERR_HIP;
        }
        j0 = j1;
    }
    if (!r) {
        const int nb = (int)std::min<int64_t>(((int64_t)n * n + 255) / 256, 65536);
        hipLaunchKernelGGL(k_sym_lower, dim3(nb), dim3(256), 0, m->st, n, h);
        (void)hipEventRecord(m->ev1, m->st);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(m->st);
        if (e != hipSuccess) { g_err = hipGetErrorString(e); r = F25_ERR_HIP; }
        float ms = 0.f;
        if (!r && hipEventElapsedTime(&ms, m->ev0, m->ev1) == hipSuccess) record_ms("gn_hessian", ms);
    }
    (void)hipStreamSynchronize(m->st);
    dfree(a);
    return r;
}

// ---- Jacobian files (pchadj.f90:340-397) --------------------------------------
// One Fortran sequential unformatted record: 4-byte length, data, length.
static bool write_record(FILE *f, const void *p, size_t bytes) {
    if (bytes > 0x7fffffffu) return false;
    const int32_t n = (int32_t)bytes;
    return fwrite(&n, 4, 1, f) == 1 && (bytes == 0 || fwrite(p, 1, bytes, f) == bytes) && fwrite(&n, 4, 1, f) == 1;
}

extern "C" int f25_write_jac(const char *path, int nrhsf, int nnpgl, int ndim, int nrec, const int *lobs,
                             const void *jac, int jac_on_device, const int *mygrad) {
    if (!path || nrhsf < 0 || nnpgl < 0 || ndim <= 0 || nrec < 0 || !lobs || !jac || !mygrad)
        FAIL(F25_ERR_ARG, "invalid arguments");
    const size_t nv = (size_t)nrhsf * nnpgl;
    std::vector<float> c(2 * nv);
    if (jac_on_device) {
        HIPCHK(hipMemcpy(c.data(), jac, nv * sizeof(float2), hipMemcpyDeviceToHost));
    } else {
        std::memcpy(c.data(), jac, nv * sizeof(float2));
    }
    std::vector<float> re(nv), im(nv);
    for (size_t i = 0; i < nv; ++i) {
        re[i] = c[2 * i];
        im[i] = c[2 * i + 1];
    }
    FILE *f = std::fopen(path, "wb");
    if (!f) FAIL(F25_ERR_ARG, "cannot open the Jacobian file");
    const int32_t hdr[4] = {nrhsf, nnpgl, ndim, nrec};
    bool ok = write_record(f, hdr, sizeof(hdr)) && write_record(f, lobs, sizeof(int32_t) * (size_t)ndim * nrec) &&
              write_record(f, re.data(), sizeof(float) * nv) && write_record(f, im.data(), sizeof(float) * nv) &&
              write_record(f, mygrad, sizeof(int32_t) * (size_t)nnpgl);
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) FAIL(F25_ERR_ARG, "writing the Jacobian file failed");
    return 0;
}

// ---- GNMFREE without allocations ----------------------------------------
// f25_gnmfree_dev_noalloc keeps everything f25_gnmfree_dev allocates per call
// (the per-Jacobian tables, the JP vectors and the column-chunk partial sums)
// in one file-scope device array.  Each GPU holds its own copy of it until
// the process ends; its address on a device comes from hipGetSymbolAddress,
// taken once per device with that device current.
static constexpr size_t kScratchBytes = size_t(32) << 20;   // 32 MiB per GPU
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
static std::mutex g_scratch_mu;
static std::map<int, unsigned char *> g_scratch_ptr;

static int scratch_on(int dev, unsigned char **out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it == g_scratch_ptr.end()) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) FAIL(F25_ERR_HIP, "hipGetSymbolAddress(scratch) failed");
        it = g_scratch_ptr.emplace(dev, static_cast<unsigned char *>(ptr)).first;
    }
    *out = it->second;
    return 0;
}

static size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

extern "C" int f25_gnmfree_dev_noalloc(f25_mesh *m, int njac, const int *nobs, const double *alpha,
                                       const void *jac, const double *p, double *z) {
    if (!m || njac <= 0 || njac > 65535 || !nobs || !alpha || !jac || !p || !z) FAIL(F25_ERR_ARG, "invalid arguments");
    if (!m->have_inv) FAIL(F25_ERR_STATE, "inversion structures missing (f25_inv_setup / f25_inv_attach)");
    HIPCHK(hipSetDevice(m->dev));
    unsigned char *base = nullptr;
    TRY(scratch_on(m->dev, &base));
    const int n = m->inv.nnpgl;
    std::vector<int64_t> joff, poff;
    int64_t jt, pt;
    TRY(gn_offsets(njac, nobs, n, joff, poff, jt, pt));
    int nch, cch;
    gn_jp_chunks(n, nch, cch);
    // scratch layout: nobs | joff | poff | alpha | JP (pt) | partials (nch * pt)
    const size_t o_nobs = 0, o_joff = align256(o_nobs + sizeof(int) * njac),
                 o_poff = align256(o_joff + sizeof(int64_t) * njac),
                 o_alpha = align256(o_poff + sizeof(int64_t) * njac),
                 o_jp = align256(o_alpha + sizeof(double) * njac), o_part = align256(o_jp + sizeof(double2) * pt),
                 need = o_part + sizeof(double2) * (size_t)nch * pt;
    if (need > kScratchBytes)
        FAIL(F25_ERR_ARG, "batch too large for the 32 MiB scratch: sum(nobs) * (chunks + 1) * 16 bytes "
                          "exceeds it; split the Jacobians over calls (z accumulates)");
    int *dn = reinterpret_cast<int *>(base + o_nobs);
    int64_t *djo = reinterpret_cast<int64_t *>(base + o_joff), *dpo = reinterpret_cast<int64_t *>(base + o_poff);
    double *da = reinterpret_cast<double *>(base + o_alpha);
    double2 *djp = reinterpret_cast<double2 *>(base + o_jp), *dpart = reinterpret_cast<double2 *>(base + o_part);
    HIPCHK(hipMemcpyAsync(dn, nobs, sizeof(int) * njac, hipMemcpyHostToDevice, m->st));
    HIPCHK(hipMemcpyAsync(djo, joff.data(), sizeof(int64_t) * njac, hipMemcpyHostToDevice, m->st));
    HIPCHK(hipMemcpyAsync(dpo, poff.data(), sizeof(int64_t) * njac, hipMemcpyHostToDevice, m->st));
    HIPCHK(hipMemcpyAsync(da, alpha, sizeof(double) * njac, hipMemcpyHostToDevice, m->st));
    (void)hipEventRecord(m->ev0, m->st);
    launch_gn_jp(m->st, n, njac, pt, dn, djo, dpo, m->d_mygrad, (const float2 *)jac, p, dpart, djp);
    hipLaunchKernelGGL(k_gn_z, dim3((n + 3) / 4), dim3(256), 0, m->st, n, njac, dn, djo, dpo, da, m->d_mygrad,
                       (const float2 *)jac, djp, z);
    (void)hipEventRecord(m->ev1, m->st);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(m->st);
    if (e != hipSuccess) FAIL(F25_ERR_HIP, hipGetErrorString(e));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, m->ev0, m->ev1) == hipSuccess) record_ms("gnmfree", ms);
    return 0;
}

extern "C" int f25_gnmfree(f25_mesh *m, int njac, const int *nobs, const double *alpha, const void *jac,
                           const double *p, double *z) {
    if (!m || njac <= 0 || njac > 65535 || !nobs || !alpha || !jac || !p || !z) FAIL(F25_ERR_ARG, "invalid arguments");
    if (!m->have_inv) FAIL(F25_ERR_STATE, "inversion structures missing (f25_inv_setup / f25_inv_attach)");
    HIPCHK(hipSetDevice(m->dev));
    const InvInfo &inv = m->inv;
    size_t tot = 0;
    for (int j = 0; j < njac; ++j) tot += (size_t)std::max(nobs[j], 0) * inv.nnpgl;
    float2 *dj = nullptr;
    double *dp_ = nullptr, *dz = nullptr;
    int r = upload(&dj, (const float2 *)jac, tot, m->st);
    if (!r) r = upload(&dp_, p, (size_t)inv.na35, m->st);
    if (!r) r = upload(&dz, z, (size_t)inv.na35, m->st);
    if (!r) r = gnmfree_core(m, njac, nobs, alpha, dj, dp_, dz);
    if (!r) {
        hipError_t e = hipMemcpyAsync(z, dz, (size_t)inv.na35 * sizeof(double), hipMemcpyDeviceToHost, m->st);
        if (e == hipSuccess) e = hipStreamSynchronize(m->st);
        if (e != hipSuccess) { g_err = hipGetErrorString(e); r = F25_ERR_HIP; }
    }
    dfree(dj); dfree(dp_); dfree(dz);
    return r;
}

// ---------------------------------------------------------------------------
// Gauss-Newton element products
extern "C" int f25_gn_fv(f25_mesh *m, int ntrip, const double *omega, const double *ky, const void *u,
                         const double *v, void *out) {
    if (m) TRY(require_elastic(m));
    if (!m || ntrip <= 0 || !u || !v || !out) FAIL(F25_ERR_ARG, "invalid arguments");
    if (!m->have_model) FAIL(F25_ERR_STATE, "f25_set_model has not been called");
    HIPCHK(hipSetDevice(m->dev));
    const int n = m->d.nnpg;
    // chain rule (Vp, Vs, rho) -> (lambda, mu, rho) perturbations
    std::vector<double> dr(n), dl(n), dm(n);
    for (int a = 0; a < n; ++a) {
        double r = m->rho[a], vp = std::sqrt((m->lam[a] + 2.0 * m->mu[a]) / r), vs = std::sqrt(m->mu[a] / r);
        // anchors of PML elements hold fixed parameters: their perturbation is ignored
        const double keep = m->fixa[a] ? 0.0 : 1.0;
        double dvp = keep * v[3 * a], dvs = keep * v[3 * a + 1], drho = keep * v[3 * a + 2];
        dr[a] = drho;
        dl[a] = drho * (vp * vp - 2.0 * vs * vs) + r * (2.0 * vp * dvp - 4.0 * vs * dvs);
        dm[a] = drho * vs * vs + 2.0 * r * vs * dvs;
    }
    double *d_r = nullptr, *d_l = nullptr, *d_m = nullptr;
    TRY(upload(&d_r, dr.data(), n, m->st));
    TRY(upload(&d_l, dl.data(), n, m->st));
    TRY(upload(&d_m, dm.data(), n, m->st));
    const size_t nb = (size_t)m->d.nelem * NBLK * m->tab.nen * m->tab.nen;
    if (!m->d_dblk) TRY(dalloc(&m->d_dblk, nb));
    int r = run_blocks(m, d_r, d_l, d_m, m->d_dblk, nullptr, nullptr, nullptr);
    dfree(d_r); dfree(d_l); dfree(d_m);
    if (r) return r;
    const size_t nn = (size_t)m->g.ndof * ntrip;
    double2 *a = nullptr, *b = nullptr;
    double *d_om = nullptr, *d_ky = nullptr;
    TRY(dalloc(&a, nn));
    TRY(dalloc(&b, nn));
    TRY(upload(&d_om, omega, ntrip, m->st));
    TRY(upload(&d_ky, ky, ntrip, m->st));
    hipError_t e = hipMemcpyAsync(a, u, nn * 16, hipMemcpyHostToDevice, m->st);
    r = (e == hipSuccess) ? 0 : F25_ERR_HIP;
    if (!r) r = transpose_dev(m->st, a, b, ntrip, m->g.ndof);
    if (!r) r = vcoef(m, m->d_dblk, &m->d_dvcf);
    if (!r) r = apply_core(m, ntrip, d_om, d_ky, F25_APPLY_F64, b, a, m->d_dvcf, nullptr);
    if (!r) r = transpose_dev(m->st, a, b, m->g.ndof, ntrip);
    if (!r) {
        e = hipMemcpyAsync(out, b, nn * 16, hipMemcpyDeviceToHost, m->st);
        if (e == hipSuccess) e = hipStreamSynchronize(m->st);
        if (e != hipSuccess) { g_err = hipGetErrorString(e); r = F25_ERR_HIP; }
    }
    dfree(a); dfree(b); dfree(d_om); dfree(d_ky);
    return r;
}

extern "C" int f25_gn_fhw(f25_mesh *m, int ntrip, const double *omega, const double *ky, const void *u,
                          const void *w, double *g) {
    return grad_host(m, ntrip, omega, ky, u, w, g, 1);
}

// patch positions of every anchor's element rows (built once per mesh)
static int ensure_pmap(f25_mesh *m) {
    if (m->d_pmap) return 0;
    const int nnpg = m->d.nnpg, nen = m->tab.nen, nee = 3 * nen, ncon = m->ncon_anchor;
    std::vector<int> cnt(nnpg, 0), pm((size_t)nnpg * ncon * nee, -1);
    std::vector<std::vector<int>> elems(nnpg);
    for (int e = 0; e < m->d.nelem; ++e)
        for (int k = 0; k < 4; ++k) elems[m->ieng[4 * e + k] - 1].push_back(e);
    std::vector<int> seen;   // DOF -> patch position, reset per anchor
    seen.assign(m->g.ndof + 1, -1);
    std::vector<int> touched;
    for (int a = 0; a < nnpg; ++a) {
        int npos = 0;
        touched.clear();
        for (size_t s = 0; s < elems[a].size(); ++s) {
            const int e = elems[a][s];
            for (int j = 0; j < nee; ++j) {
                const int dof = m->g.lm[(size_t)e * nee + j];   // (nen, 3) rows of e, 1-based
                if (dof <= 0) continue;
                if (seen[dof] < 0) { seen[dof] = npos++; touched.push_back(dof); }
                pm[((size_t)a * ncon + s) * nee + j] = seen[dof];
            }
        }
        for (int d : touched) seen[d] = -1;
    }
    // element order per anchor must match the device connectivity (mconn)
    return upload(&m->d_pmap, pm.data(), pm.size(), m->st);
}

extern "C" int f25_gn_diag(f25_mesh *m, int ntrip, const double *omega, const double *ky, const void *u,
                           double *diag) {
    if (m) TRY(require_elastic(m));
    if (!m || ntrip <= 0 || !omega || !ky || !u || !diag) FAIL(F25_ERR_ARG, "invalid arguments");
    if (!m->have_model) FAIL(F25_ERR_STATE, "f25_set_model has not been called");
    HIPCHK(hipSetDevice(m->dev));
    TRY(ensure_pmap(m));
    const int nen = m->tab.nen, nee = 3 * nen, nq = m->tab.nq, nnpg = m->d.nnpg;
    const size_t ndof = m->g.ndof;
    double2 *du = nullptr, *dcols = nullptr;
    double *dd = nullptr;
    int r = dalloc(&du, ndof * ntrip);
    if (!r) r = dalloc(&dcols, (size_t)m->d.nelem * 12 * nee);
    if (!r) r = upload(&dd, diag, (size_t)nnpg * 3, m->st);
    if (!r) {
        hipError_t e = hipMemcpyAsync(du, u, ndof * ntrip * 16, hipMemcpyHostToDevice, m->st);
        if (e != hipSuccess) { g_err = hipGetErrorString(e); r = F25_ERR_HIP; }
    }
    // the sum-factorised element kernel (every order up to p6)
    typedef void (*FcolsSf)(const double *, const double *, const double *, const int *, const double *,
                            const double *, const char *, const int *, const double *, const double *,
                            const double *, double, double, const double2 *, double2 *, int *);
    FcolsSf ksf = nullptr;
    switch (m->tab.nl) {
        case 3: ksf = k_fcols_sf<3>; break;
        case 4: ksf = k_fcols_sf<4>; break;
        case 5: ksf = k_fcols_sf<5>; break;
        case 6: ksf = k_fcols_sf<6>; break;
        case 7: ksf = k_fcols_sf<7>; break;
        default: break;
    }
    if (!ksf && !r) {
        g_err = "unsupported order";
        r = F25_ERR_ARG;
    }
    const int nl = m->tab.nl, q1 = m->tab.nq1;
    // Q1 = 17 (p6): five waves and the 17 point fields of the xi = 16 column
    const int ntsf = q1 > 16 ? 320 : 256;
    const size_t lsf = (size_t)(2 * nl * q1 + 2 * q1) * sizeof(double) +
                       (size_t)std::max(nee + 6 * q1 * nl, 28 * q1 * nl) * sizeof(double2) +
                       (q1 > 16 ? (size_t)17 * q1 * sizeof(double2) + (size_t)2 * nl * q1 * sizeof(double) : 0);
    const int nt2 = 256;
    const size_t lds2 = (size_t)m->ncon_anchor * nee * 3 * sizeof(double2);
    for (int k = 0; k < 2 && !r; ++k) {
        const size_t l = k ? lds2 : lsf;
        if (l <= 65536) continue;
        hipError_t e = hipFuncSetAttribute(k ? (const void *)k_fdiag : (const void *)ksf,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)l);
        if (e != hipSuccess) { g_err = hipGetErrorString(e); r = F25_ERR_HIP; }
    }
    if (!r) (void)hipEventRecord(m->ev0, m->st);
    for (int t = 0; t < ntrip && !r; ++t) {
        hipLaunchKernelGGL(ksf, dim3(m->d.nelem), dim3(ntsf), lsf, m->st, m->d_t1d, m->d_wq, m->d_xq, m->d_ieng,
                           m->d_xl, m->d_zl, m->d_cdom, m->d_lm, m->d_rho, m->d_lam, m->d_mu, omega[t], ky[t],
                           du + ndof * t, dcols, m->d_err);
        hipLaunchKernelGGL(k_fdiag, dim3(nnpg), dim3(nt2), lds2, m->st, nnpg, m->ncon_anchor, nee,
                           m->d_mconn, m->d_ieng, m->d_pmap, m->d_fixa, dcols, dd);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) { g_err = hipGetErrorString(e); r = F25_ERR_HIP; }
    }
    if (!r) (void)hipEventRecord(m->ev1, m->st);
    if (!r) r = check_err(m);
    float dms = 0.f;
    if (!r && hipEventElapsedTime(&dms, m->ev0, m->ev1) == hipSuccess) record_ms("gn_diag", dms);
    if (!r) {
        hipError_t e = hipMemcpyAsync(diag, dd, (size_t)nnpg * 3 * sizeof(double), hipMemcpyDeviceToHost, m->st);
        if (e == hipSuccess) e = hipStreamSynchronize(m->st);
        if (e != hipSuccess) { g_err = hipGetErrorString(e); r = F25_ERR_HIP; }
    }
    dfree(du); dfree(dcols); dfree(dd);
    return r;
}

// Workgroup residency: the most workgroups of nt threads, holding vgprs
// registers per lane and lds bytes of LDS, resident at once on one CU
// (median over the CUs of the per-CU maximum), from the realtime stamps of
// 8 workgroups per CU that each spin ~40,000 cycles.
extern "C" int f25_selftest_residency(int nt, int vgprs, int lds, int *groups_per_cu) {
    if (!groups_per_cu) FAIL(F25_ERR_ARG, "null result pointer");
    if (nt != 192 && nt != 256 && nt != 320 && nt != 384 && nt != 512)
        FAIL(F25_ERR_ARG, "threads per workgroup: 192, 256, 320, 384 or 512");
    if (vgprs != 128 && vgprs != 168) FAIL(F25_ERR_ARG, "registers per lane: 128 or 168");
    if (lds < 0 || lds > 160 * 1024) FAIL(F25_ERR_ARG, "LDS bytes: 0 .. 160 KiB");
    int dev = 0, ncu = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const void *k = nullptr;
#define F25_PICK(NT)                                                                        \
    if (nt == NT) k = vgprs == 128 ? (const void *)k_resid_probe_128<NT> : (const void *)k_resid_probe_168<NT>;
    F25_PICK(192) F25_PICK(256) F25_PICK(320) F25_PICK(384) F25_PICK(512)
#undef F25_PICK
    const int nwg = 8 * ncu;
    unsigned long long *d = nullptr;
    TRY(dalloc(&d, (size_t)3 * nwg));
    hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e == hipSuccess) e = hipMemset(d, 0, (size_t)3 * nwg * sizeof(unsigned long long));
    unsigned long long spin = 40000;
    int nw = nwg;
    void *args[] = {&d, &nw, &spin};
    if (e == hipSuccess) e = hipLaunchKernel(k, dim3(nwg), dim3(nt), args, (size_t)lds, 0);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    std::vector<unsigned long long> h((size_t)3 * nwg);
    if (e == hipSuccess) e = hipMemcpy(h.data(), d, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost);
    dfree(d);
    if (e != hipSuccess) FAIL(F25_ERR_HIP, hipGetErrorString(e));
    // per CU (XCC_ID, then HW_ID bits 8..15: CU, shader array, engine): the
    // largest number of workgroups alive at one instant
    std::map<unsigned long long, std::vector<std::pair<unsigned long long, int>>> ev;
    for (int w = 0; w < nwg; ++w) {
        const unsigned long long *o = &h[(size_t)3 * w];
        if (!o[0] || o[1] < o[0]) FAIL(F25_ERR_HIP, "residency probe: a workgroup left no record");
        const unsigned long long cu = ((o[2] >> 32) & 0xf) << 8 | ((o[2] >> 8) & 0xff);
        ev[cu].push_back({o[0], 1});
        ev[cu].push_back({o[1], -1});
    }
    std::vector<int> mx;
    for (auto &kv : ev) {
        auto &v = kv.second;
        std::sort(v.begin(), v.end());   // at equal stamps an end (-1) sorts first
        int cur = 0, m = 0;
        for (auto &x : v) { cur += x.second; m = std::max(m, cur); }
        mx.push_back(m);
    }
    std::sort(mx.begin(), mx.end());
    *groups_per_cu = mx[mx.size() / 2];
    return 0;
}

// MFMA operand/result layout check (used by the self test entry point)
extern "C" int f25_selftest_mfma(double *maxerr) {
    double *dd = nullptr;
    float *df = nullptr;
    TRY(dalloc(&dd, 256));
    TRY(dalloc(&df, 256));
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, 0, dd, df);
    std::vector<double> hd(256);
    std::vector<float> hf(256);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(hd.data(), dd, 256 * 8, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(hf.data(), df, 256 * 4, hipMemcpyDeviceToHost);
    dfree(dd);
    dfree(df);
    if (e != hipSuccess) FAIL(F25_ERR_HIP, hipGetErrorString(e));
    // A[row][k] = 1 + row + 16k (lane row + 16k), B[k][col] = 1 + 0.5 col - 3k
    double worst = 0.0;
    auto ref = [](int row, int col) {
        double r = 0.0;
        for (int k = 0; k < 4; ++k) r += (1.0 + row + 16.0 * k) * (1.0 + 0.5 * col - 3.0 * k);
        return r;
    };
    for (int l = 0; l < 64; ++l)
        for (int i = 0; i < 4; ++i) {
            const int col = l & 15;
            worst = std::max(worst, std::fabs(hd[l * 4 + i] - ref((l >> 4) + 4 * i, col)));
            worst = std::max(worst, std::fabs((double)hf[l * 4 + i] - ref(4 * (l >> 4) + i, col)));
        }
    if (maxerr) *maxerr = worst;
    return worst == 0.0 ? 0 : F25_ERR_HIP;
}
