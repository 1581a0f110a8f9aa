// HIP kernels for gfx950 (MI355X): element blocks, batched element-block
// product (FP64 on the vector ALUs, FP32 on MFMA), CSR assembly, PML
// elements, SpMV, adjoint gradient, the reference virtual-force columns and
// their CSC gather.
//
// Element operator structure (estiff25.f:61-150, restated as bilinear
// forms).  For a non-PML element with interpolated lambda, mu, rho and
// quadrature weight w = wx*wz*det J, the element matrix of KM25ISOQ is
//   y1 = K11 x1 + K13 x3 - i ky T12 x2 + (ky^2 Mmu - w^2 Mrho) x1
//   y2 = K22 x2 + i ky T12^T x1 + i ky T23 x3 + (ky^2 Ml2m - w^2 Mrho) x2
//   y3 = K13^T x1 + K33 x3 - i ky T23^T x2 + (ky^2 Mmu - w^2 Mrho) x3
// with nine real NEN x NEN blocks that do not depend on omega or ky:
//   K11 = S[(l+2m) Nx Nx + m Nz Nz]   K22 = S[m (Nx Nx + Nz Nz)]
//   K33 = S[(l+2m) Nz Nz + m Nx Nx]   K13 = S[l Nx_a Nz_b + m Nz_a Nx_b]
//   T12 = S[l Nx_a N_b - m N_a Nx_b]  T23 = S[l N_a Nz_b - m Nz_a N_b]
//   Mrho = S[rho N N], Mmu = S[m N N], Ml2m = S[(l+2m) N N]
// (S = sum over quadrature points with weight w).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>
#include <utility>

namespace f25k {

enum { B_K11 = 0, B_K22, B_K33, B_K13, B_T12, B_T23, B_MRHO, B_MMU, B_ML2M, NBLK };

struct Geo {   // geometry at one quadrature point
    double det, dxdxi, dxdeta, dzdxi, dzdeta;
    double sf[4];
};

__device__ inline int geo_at(double xi, double eta, const int *ie, const double *xl,
                             const double *zl, Geo &g) {
    const double eps = 2.2204e-16;
    double sf[4] = {.25 * (1.0 - xi) * (1.0 - eta), .25 * (1.0 + xi) * (1.0 - eta),
                    .25 * (1.0 + xi) * (1.0 + eta), .25 * (1.0 - xi) * (1.0 + eta)};
    double dxi[4] = {-.25 * (1.0 - eta), .25 * (1.0 - eta), .25 * (1.0 + eta), -.25 * (1.0 + eta)};
    double deta[4] = {-.25 * (1.0 - xi), -.25 * (1.0 + xi), .25 * (1.0 + xi), .25 * (1.0 - xi)};
    double a = 0, b = 0, c = 0, d = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (fabs(sf[k]) < eps) sf[k] = 0.0;
        if (fabs(dxi[k]) < eps) dxi[k] = 0.0;
        if (fabs(deta[k]) < eps) deta[k] = 0.0;
        int gi = ie[k] - 1;
        a += dxi[k] * xl[gi];
        b += deta[k] * xl[gi];
        c += dxi[k] * zl[gi];
        d += deta[k] * zl[gi];
        g.sf[k] = sf[k];
    }
    g.dxdxi = a; g.dxdeta = b; g.dzdxi = c; g.dzdeta = d;
    g.det = a * d - b * c;
    return (fabs(g.det) <= 2.22e-16 || g.det < 0.0) ? 1 : 0;   // cjac.f:115-125
}

// ---------------------------------------------------------------------------
// Element blocks.  One workgroup per element; every thread owns up to PPT
// (a, b) node pairs and accumulates the nine blocks over quadrature points
// staged in LDS chunk by chunk.
template <int NL, int NT>
__global__ __launch_bounds__(NT) void k_blocks(int nelem, int nq1, const double *__restrict__ shl,
                                               const double *__restrict__ wq,
                                               const double *__restrict__ xq,
                                               const int *__restrict__ ieng,
                                               const double *__restrict__ xl,
                                               const double *__restrict__ zl,
                                               const double *__restrict__ rho,
                                               const double *__restrict__ lam,
                                               const double *__restrict__ mu,
                                               const char *__restrict__ cdom, double *__restrict__ blk,
                                               float *__restrict__ blk32, float *__restrict__ blk_hi,
                                               float *__restrict__ blk_lo, int *__restrict__ err) {
    constexpr int NEN = NL * NL;
    constexpr int NP = NEN * NEN;
    constexpr int PPT = (NP + NT - 1) / NT;
    constexpr int QC = 16;  // quadrature points per LDS chunk
    __shared__ double sN[QC][NEN], sX[QC][NEN], sZ[QC][NEN];
    __shared__ double sW[QC][4];   // w*rho, w*lambda, w*mu
    const int e = blockIdx.x;
    if (e >= nelem) return;
    const int tid = threadIdx.x;
    const size_t obase = (size_t)e * NBLK * NP;
    if (cdom[e] == 'A') {   // PML elements are integrated per call
        for (int i = tid; i < NBLK * NP; i += NT) {
            blk[obase + i] = 0.0;
            if (blk32) blk32[obase + i] = 0.f;
            if (blk_hi) { blk_hi[obase + i] = 0.f; blk_lo[obase + i] = 0.f; }
        }
        return;
    }
    const int *ie = ieng + 4 * e;
    double acc[PPT][NBLK];
#pragma unroll
    for (int p = 0; p < PPT; ++p)
#pragma unroll
        for (int k = 0; k < NBLK; ++k) acc[p][k] = 0.0;
    const int nq = nq1 * nq1;
    for (int q0 = 0; q0 < nq; q0 += QC) {
        __syncthreads();
        for (int i = tid; i < QC * NEN; i += NT) {
            int qq = i / NEN, a = i % NEN, q = q0 + qq;
            if (q < nq) {
                int ix = q % nq1, iz = q / nq1;
                Geo g;
                int bad = geo_at(xq[ix], xq[iz], ie, xl, zl, g);
                if (bad && a == 0) *err = 2;
                const double *s = shl + ((size_t)q * NEN + a) * 3;
                double rj = bad ? 1.0 : g.det;
                sN[qq][a] = s[2];
                sX[qq][a] = (s[0] * g.dzdeta - s[1] * g.dzdxi) / rj;
                sZ[qq][a] = -(s[0] * g.dxdeta - s[1] * g.dxdxi) / rj;
                if (a == 0) {
                    double w = wq[ix] * wq[iz] * g.det, r = 0, l = 0, m = 0;
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        int gi = ie[k] - 1;
                        r += g.sf[k] * rho[gi];
                        l += g.sf[k] * lam[gi];
                        m += g.sf[k] * mu[gi];
                    }
                    sW[qq][0] = w * r; sW[qq][1] = w * l; sW[qq][2] = w * m;
                }
            } else {
                sN[qq][a] = sX[qq][a] = sZ[qq][a] = 0.0;
                if (a == 0) sW[qq][0] = sW[qq][1] = sW[qq][2] = 0.0;
            }
        }
        __syncthreads();
#pragma unroll
        for (int p = 0; p < PPT; ++p) {
            int pr = tid + p * NT;
            if (pr >= NP) break;
            int a = pr / NEN, b = pr % NEN;
#pragma unroll 2
            for (int qq = 0; qq < QC; ++qq) {
                double wr = sW[qq][0], wl = sW[qq][1], wm = sW[qq][2];
                double na = sN[qq][a], nb = sN[qq][b], xa = sX[qq][a], xb = sX[qq][b];
                double za = sZ[qq][a], zb = sZ[qq][b];
                double nn = na * nb, xx = xa * xb, zz = za * zb;
                acc[p][B_K11] += (wl + 2.0 * wm) * xx + wm * zz;
                acc[p][B_K22] += wm * (xx + zz);
                acc[p][B_K33] += (wl + 2.0 * wm) * zz + wm * xx;
                acc[p][B_K13] += wl * xa * zb + wm * za * xb;
                acc[p][B_T12] += wl * xa * nb - wm * na * xb;
                acc[p][B_T23] += wl * na * zb - wm * za * nb;
                acc[p][B_MRHO] += wr * nn;
                acc[p][B_MMU] += wm * nn;
                acc[p][B_ML2M] += (wl + 2.0 * wm) * nn;
            }
        }
    }
    // Split for the refined FP32 product: hi = v rounded to a multiple of
    // 2^(E-11), where 2^E bounds the whole block, so |hi| / 2^(E-11) <= 2^11
    // (12 significant bits, exact in FP32); lo = fl32(v - hi).
    __shared__ int sExp[NBLK];
    if (blk_hi) {
        __shared__ double sMax[NT / 64][NBLK];
        double mx[NBLK];
#pragma unroll
        for (int k = 0; k < NBLK; ++k) mx[k] = 0.0;
#pragma unroll
        for (int p = 0; p < PPT; ++p) {
            int pr = tid + p * NT;
            if (pr >= NP) break;
#pragma unroll
            for (int k = 0; k < NBLK; ++k) mx[k] = fmax(mx[k], fabs(acc[p][k]));
        }
#pragma unroll
        for (int k = 0; k < NBLK; ++k)
            for (int off = 32; off > 0; off >>= 1) mx[k] = fmax(mx[k], __shfl_xor(mx[k], off));
        if ((tid & 63) == 0)
#pragma unroll
            for (int k = 0; k < NBLK; ++k) sMax[tid >> 6][k] = mx[k];
        __syncthreads();
        if (tid < NBLK) {
            double v = 0.0;
            for (int w = 0; w < NT / 64; ++w) v = fmax(v, sMax[w][tid]);
            int ex = 0;
            (void)frexp(v, &ex);   // v < 2^ex
            sExp[tid] = ex;
        }
        __syncthreads();
    }
#pragma unroll
    for (int p = 0; p < PPT; ++p) {
        int pr = tid + p * NT;
        if (pr >= NP) break;
#pragma unroll
        for (int k = 0; k < NBLK; ++k) {
            double v = acc[p][k];
            blk[obase + (size_t)k * NP + pr] = v;
            if (blk32) blk32[obase + (size_t)k * NP + pr] = (float)v;
            if (blk_hi) {
                const int ex = sExp[k];
                const double h = ldexp(rint(ldexp(v, 11 - ex)), ex - 11);
                blk_hi[obase + (size_t)k * NP + pr] = (float)h;
                blk_lo[obase + (size_t)k * NP + pr] = (float)(v - h);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Batched element-block product on MFMA: the FP32 and refined FP32 modes
// (FP32 MFMA runs at twice the rate of FP32 FMAs on the vector ALUs).  The
// FP64 product is k_apply_v below.
//
// One workgroup = one element of the current colour.  The element's blocks
// sit in LDS (row stride RS, odd, so both the direct and the transposed
// operand reads are conflict-free).  Each wave owns N-tiles of 8 (omega, ky)
// pairs = 16 real columns (re, im interleaved as in memory).  The input
// vector values reach the B operand of a 16x16x4 MFMA directly (lane l:
// column l%16, k-row l/16 + 4s); per-column factors (ky, ky^2, -omega^2, the
// i of the coupling blocks) are applied to the B operand, so each output
// component is ONE accumulation over five sub-blocks.  Outputs are added
// into y[dof][pair] (no other element of the colour touches these DOFs).
template <typename T>
struct MfmaTraits;
// Operand layout of both 16x16x4 forms: lane l supplies A[l%16][l/16] and
// B[l/16][l%16].  Result element i (of 4) of lane l is column l%16 and row
//   f64: l/16 + 4 i        f32: 4 (l/16) + i
// (measured on gfx950 with a single-wave probe).
template <>
struct MfmaTraits<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    __device__ static inline v4 mma(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static constexpr int row(int kg, int i) { return kg + 4 * i; }
};
template <>
struct MfmaTraits<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    __device__ static inline v4 mma(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static constexpr int row(int kg, int i) { return 4 * kg + i; }
};

struct SubBlk { int blk; int trans; int xcomp; int kind; };
// kind: 0 -> 1, 1 -> +i ky, 2 -> -i ky, 3 -> ky^2, 4 -> -omega^2
__device__ constexpr SubBlk kSub[3][5] = {
    {{B_K11, 0, 0, 0}, {B_K13, 0, 2, 0}, {B_T12, 0, 1, 2}, {B_MMU, 0, 0, 3}, {B_MRHO, 0, 0, 4}},
    {{B_K22, 0, 1, 0}, {B_T12, 1, 0, 1}, {B_T23, 0, 2, 1}, {B_ML2M, 0, 1, 3}, {B_MRHO, 0, 1, 4}},
    {{B_K13, 1, 0, 0}, {B_K33, 0, 2, 0}, {B_T23, 1, 1, 2}, {B_MMU, 0, 2, 3}, {B_MRHO, 0, 2, 4}}};

template <int NL>
struct ApplyShape {
    static constexpr int NEN = NL * NL;
    static constexpr int RS = (NEN % 2) ? NEN : NEN + 1;
    static constexpr int MT = (NEN + 15) / 16;
    static constexpr int KS = (NEN + 3) / 4;
    // a last row tile holding at most 2 element rows (p6: 1) is done on the
    // vector ALUs instead of a mostly empty 16-row MFMA tile (with 4 rows,
    // p5, the extra LDS reads cost more than the MFMA tile saves)
    static constexpr int REM = NEN - 16 * (MT - 1);
    static constexpr int NR = (MT > 1 && REM <= 2) ? REM : 0;   // rows on VALU
    static constexpr int MTM = NR ? MT - 1 : MT;                 // row tiles on MFMA
    // NEN = 4 KS - 3 (p2, p4, p6): the last k step holds one node.  Its five
    // sub-block products into one accumulator are merged: the k index of the
    // MFMA runs over sub-blocks 0..3 (A column NEN-1 of each, their operands
    // of node NEN-1), and sub-block 4 takes a second MFMA: 2 instead of 5.
    static constexpr bool KM = (NEN % 4 == 1);
};

// LDS bytes of k_apply for a shape.  PH: the nine blocks do not fit, so the
// product runs as three launches, one output component each, holding that
// component's five blocks.  PRE: the tile's y values are loaded at the start
// of the tile, next to the first x operands (one memory latency per tile).
// NW: waves per workgroup, chosen so that a CU (160 KB of LDS) holds about
// twelve waves whatever the footprint: 4 waves when three elements fit, 12
// when only one does.
template <int NL, typename BT, bool REFINE>
struct ApplyLds {
    typedef ApplyShape<NL> S;
    static constexpr size_t per_blk = (size_t)S::NEN * S::RS * sizeof(BT) * (REFINE ? 2 : 1);
    static constexpr bool PH = 9 * per_blk > 150 * 1024;
    static constexpr size_t bytes = (PH ? 5 : 9) * per_blk + 3 * S::NEN * sizeof(int);
    // (the refined product runs at two waves per SIMD and measured 3% faster
    // without it: 7.71 -> 7.49 ms per step at 200x100, p4)
    static constexpr bool PRE = S::MT <= 2 && !REFINE;
    static constexpr int per_cu = (int)((160 * 1024) / bytes);
    // (the refined product needs more registers: at most 8 waves per element)
    static constexpr int NW = per_cu >= 3 ? 4 : (per_cu == 2 ? 6 : (REFINE ? 8 : 12));
};

// value held by lane l^1 (the re/im partner of a column pair), quad_perm [1,0,3,2]
__device__ inline double partner(double v) {
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_update_dpp(0, lo, 0xB1, 0xF, 0xF, false);
    hi = __builtin_amdgcn_update_dpp(0, hi, 0xB1, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
__device__ inline float partner(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
}

// XT: storage type of x/y (double2 or float2); BT: block/MFMA type.
// CSEL: -1 all three output components, 0..2 one component (PH launches).
// REFINE: blocks and the double operand each split into a fixed-point hi
// part (12 bits on a per-block / per-pair-and-kind exponent grid, so the
// four-term hi*hi partial sums of one MFMA are exact in FP32) and an FP32
// remainder; hi*hi, hi*lo, lo*hi and lo*lo on FP32 MFMA, summed in double.
// first[e] bit a: element e is the first (lowest colour) to touch node a, so
// it writes y instead of adding to it (no zero fill of y is needed).
//
// Loop order per tile: k-steps s (4 element nodes each) outermost, with the
// x operands of step s+1 loaded while step s runs; inside a step every
// (component, sub-block, row tile) accumulator gets its MFMA, so the five
// operand kinds of one x value are formed once and reused.  At least three
// waves per SIMD (two for the refined product) are requested: the register
// budget that leaves (168) keeps the y preload and costs a few spilled
// values, which measured faster than either two waves or no preload.
template <int NL, typename BT, typename XT, bool REFINE, int CSEL>
__global__ __launch_bounds__((64 * ApplyLds<NL, BT, REFINE>::NW), (REFINE ? 2 : 3)) void k_apply(const int *__restrict__ order,
                                               const BT *__restrict__ blk, const BT *__restrict__ blk_lo,
                                               const int *__restrict__ lm,
                                               const unsigned long long *__restrict__ first, int npair,
                                               const double *__restrict__ omega,
                                               const double *__restrict__ ky,
                                               const double *__restrict__ colmax,
                                               const XT *__restrict__ x, XT *__restrict__ y) {
    typedef ApplyShape<NL> S;
    typedef ApplyLds<NL, BT, REFINE> LD;
    constexpr int NEN = S::NEN, RS = S::RS, MT = S::MTM, KS = S::KS, NR = S::NR;
    // merged last k step: with the three-component shapes (p2, p4) the
    // extra selects push the 168-register budget into spills that cost more
    // than the three MFMAs saved (p4: 3.15 -> 3.35 ms per step); the
    // one-component launches of p6 have the room (66.5 -> 65.1 ms)
    constexpr bool KM = S::KM && !REFINE && (CSEL >= 0);
    constexpr int NP = NEN * NEN;
    constexpr bool PRE = LD::PRE;
    constexpr int NC = (CSEL < 0) ? 3 : 1;
    constexpr int NSLOT = (CSEL < 0) ? 9 : 5;
    typedef MfmaTraits<BT> MT_;
    typedef typename MT_::v4 v4;
    typedef decltype(XT().x) XS;                                        // float or double
    typedef typename std::conditional<REFINE, double, XS>::type OS;     // operand arithmetic
    extern __shared__ __align__(16) unsigned char smem[];
    BT *sB = reinterpret_cast<BT *>(smem);
    BT *sBl = sB + (REFINE ? NSLOT * NEN * RS : 0);
    int *sLM = reinterpret_cast<int *>(sB + (REFINE ? 2 : 1) * NSLOT * NEN * RS);
    const int e = order[blockIdx.x];
    const int tid = threadIdx.x;
    const size_t gbase = (size_t)e * NBLK * NP;
    const unsigned long long fmask = first[e];
    constexpr int NW = LD::NW, NTH = 64 * NW;
    for (int i = tid; i < 3 * NEN; i += NTH) sLM[i] = lm[(size_t)e * 3 * NEN + i];
    for (int i = tid; i < NSLOT * NP; i += NTH) {
        const int sl = i / NP, r = i % NP, a = r / NEN, b = r % NEN;
        const int kb = (CSEL < 0) ? sl : kSub[CSEL < 0 ? 0 : CSEL][sl].blk;
        sB[(sl * NEN + a) * RS + b] = blk[gbase + (size_t)kb * NP + r];
        if (REFINE) sBl[(sl * NEN + a) * RS + b] = blk_lo[gbase + (size_t)kb * NP + r];
    }
    __syncthreads();
    const int lane = tid & 63, wave = tid >> 6;
    const int col = lane & 15, kg = lane >> 4, ri = col & 1;
    const int ntile = (npair + 7) / 8;
    // x operands of node 4s + kg for tile tt (zero outside the element, for
    // fixed DOFs and for pairs beyond npair)
    auto load_x = [&](int tt, int s, OS *dst) {
        const int pr = tt * 8 + (col >> 1);
        const int b = (KM && s == KS - 1) ? NEN - 1 : 4 * s + kg;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int dof = (b < NEN && pr < npair) ? sLM[b * 3 + d] : -1;
            dst[d] = dof >= 0 ? (OS) reinterpret_cast<const XS *>(x + (size_t)dof * npair + pr)[ri] : (OS)0;
        }
    };
    // the first step of the next tile is loaded during the last step of the
    // current one
    OS xn[3];
    if (wave < ntile) load_x(wave, 0, xn);
#pragma unroll 1
    for (int t = wave; t < ntile; t += NW) {
        const int pair = t * 8 + (col >> 1);
        const bool live = pair < npair;
        const OS om = live ? (OS)omega[pair] : (OS)0, kv = live ? (OS)ky[pair] : (OS)0;
        const OS f1 = ri ? kv : -kv, f3 = kv * kv, f4 = -om * om;
        // refined mode: per operand kind, 2^eB bounds |kind factor| * max|x| of
        // this pair, so the hi part of an operand is a multiple of 2^(eB-11)
        int eB[5] = {0, 0, 0, 0, 0};
        if (REFINE) {
            const double cm = live ? colmax[pair] : 0.0;
            const double kap[5] = {1.0, fabs((double)kv), fabs((double)kv), (double)(kv * kv), (double)(om * om)};
#pragma unroll
            for (int j = 0; j < 5; ++j) (void)frexp(cm * kap[j], &eB[j]);
        }
        // the y preload is branch-free: rows outside the element, fixed DOFs
        // and first touches read a valid address (row 0) and are zeroed
        const int pcl = live ? pair : 0;
        XS yv[PRE ? NC : 1][PRE ? MT : 1][4];
        if (PRE) {
#pragma unroll
            for (int cc = 0; cc < NC; ++cc)
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int c = (CSEL < 0) ? cc : CSEL;
                        const int row = mt * 16 + MT_::row(kg, i);
                        const int dof = sLM[min(row, NEN - 1) * 3 + c];
                        const bool add = row < NEN && live && dof >= 0 && !((fmask >> row) & 1ull);
                        const XS v = reinterpret_cast<const XS *>(y + (size_t)(add ? dof : 0) * npair + pcl)[ri];
                        yv[PRE ? cc : 0][PRE ? mt : 0][i] = add ? v : (XS)0;
                    }
        }
        v4 acc[NC][MT];
        double accd[REFINE ? NC : 1][REFINE ? MT : 1][4];
        // rows 16*MT + r (r < NR): this lane's partial sum over its k values
        OS rem[NC][NR > 0 ? NR : 1];
#pragma unroll
        for (int cc = 0; cc < NC; ++cc) {
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                acc[cc][mt] = v4{0, 0, 0, 0};
#pragma unroll
                for (int i = 0; i < 4; ++i) accd[REFINE ? cc : 0][REFINE ? mt : 0][i] = 0.0;
            }
#pragma unroll
            for (int r = 0; r < (NR > 0 ? NR : 1); ++r) rem[cc][r] = (OS)0;
        }
#pragma unroll 1
        for (int s = 0; s < (KM ? KS - 1 : KS); ++s) {
            const OS xv[3] = {xn[0], xn[1], xn[2]};
            if (s + 1 < KS) load_x(t, s + 1, xn);
            else if (t + NW < ntile) load_x(t + NW, 0, xn);
            const int k = min(4 * s + kg, NEN - 1);
            // operand kinds of each x component: 0 x, 1 +i ky x, 2 -i ky x, 3 ky^2 x, 4 -w^2 x
            OS bop[3][5];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const OS p = partner(xv[d]);
                bop[d][0] = xv[d];
                bop[d][1] = f1 * p;
                bop[d][2] = -f1 * p;
                bop[d][3] = f3 * xv[d];
                bop[d][4] = f4 * xv[d];
            }
            // refined mode: fixed-point hi (12 bits, exact in FP32) + FP32 rest
            float fh[REFINE ? 3 : 1][REFINE ? 5 : 1], fl[REFINE ? 3 : 1][REFINE ? 5 : 1];
            if (REFINE) {
#pragma unroll
                for (int d = 0; d < 3; ++d)
#pragma unroll
                    for (int j = 0; j < 5; ++j) {
                        const double v = (double)bop[d][j];
                        const double h = ldexp(rint(ldexp(v, 11 - eB[j])), eB[j] - 11);
                        fh[REFINE ? d : 0][REFINE ? j : 0] = (float)h;
                        fl[REFINE ? d : 0][REFINE ? j : 0] = (float)(v - h);
                    }
            }
#pragma unroll
            for (int cc = 0; cc < NC; ++cc) {
                constexpr int C0 = (CSEL < 0) ? 0 : CSEL;
#pragma unroll
                for (int j = 0; j < 5; ++j) {
                    const SubBlk sb = kSub[(CSEL < 0) ? cc : C0][j];
                    const int slot = (CSEL < 0) ? sb.blk : j;
                    const BT *A = sB + slot * NEN * RS;
                    const BT *Al = sBl + slot * NEN * RS;
                    const OS bv = bop[sb.xcomp][sb.kind];
#pragma unroll
                    for (int mt = 0; mt < MT; ++mt) {
                        // rows >= NEN give result rows that are never stored and the
                        // operand is zero for k >= NEN: indices are clamped, not masked
                        const int arow = min(mt * 16 + col, NEN - 1);
                        const int idx = sb.trans ? (k * RS + arow) : (arow * RS + k);
                        if (REFINE) {
                            // hi*hi: products of 12-bit multiples of 2^(eA-11) and
                            // 2^(eB-11), four of them summed: exact in FP32
                            const float bh = fh[REFINE ? sb.xcomp : 0][REFINE ? sb.kind : 0];
                            const float bl = fl[REFINE ? sb.xcomp : 0][REFINE ? sb.kind : 0];
                            const BT av = A[idx], avl = Al[idx];
                            const v4 z0 = {0, 0, 0, 0};
                            const v4 z1 = MT_::mma(av, bh, z0);
                            v4 z2 = MT_::mma(av, bl, z0);
                            z2 = MT_::mma(avl, bh, z2);
                            z2 = MT_::mma(avl, bl, z2);
#pragma unroll
                            for (int i = 0; i < 4; ++i)
                                accd[REFINE ? cc : 0][REFINE ? mt : 0][i] += (double)z1[i] + (double)z2[i];
                        } else {
                            acc[cc][mt] = MT_::mma(A[idx], (BT)bv, acc[cc][mt]);
                        }
                    }
#pragma unroll
                    for (int r = 0; r < NR; ++r) {
                        const int row = MT * 16 + r;
                        const int idx = sb.trans ? (k * RS + row) : (row * RS + k);
                        const OS av = REFINE ? (OS)((double)A[idx] + (double)Al[idx]) : (OS)A[idx];
                        rem[cc][r] += av * bv;
                    }
                }
            }
        }
        if (KM) {
            // merged last step: every lane holds node NEN-1 of its column;
            // lane group kg supplies sub-block j = kg (A column NEN-1 and that
            // sub-block's operand: fa x + fb partner(x), exact for the five
            // kinds), then sub-block 4 on the kg = 0 lanes
            constexpr int kn = NEN - 1;
            const OS xv[3] = {xn[0], xn[1], xn[2]};
            if (t + NW < ntile) load_x(t + NW, 0, xn);
#pragma unroll
            for (int cc = 0; cc < NC; ++cc) {
                const int c = (CSEL < 0) ? cc : CSEL;
                const SubBlk s0 = kSub[c][0], s1 = kSub[c][1], s2 = kSub[c][2], s3 = kSub[c][3],
                             s4 = kSub[c][4];
                const SubBlk sj = kg == 0 ? s0 : kg == 1 ? s1 : kg == 2 ? s2 : s3;
                const OS xj = sj.xcomp == 0 ? xv[0] : sj.xcomp == 1 ? xv[1] : xv[2];
                const OS pj = partner(xj);
                const OS fa = sj.kind == 0 ? (OS)1 : sj.kind == 3 ? f3 : sj.kind == 4 ? f4 : (OS)0;
                const OS fb = sj.kind == 1 ? f1 : sj.kind == 2 ? -f1 : (OS)0;
                const OS bj = fa * xj + fb * pj;
                const OS b4 = kg == 0 ? f4 * xv[s4.xcomp] : (OS)0;   // sub-block 4: -w^2 x
                const BT *Aj = sB + ((CSEL < 0) ? sj.blk : kg) * NEN * RS;
                const BT *A4 = sB + ((CSEL < 0) ? s4.blk : 4) * NEN * RS;
#pragma unroll
                for (int mt = 0; mt < MT; ++mt) {
                    const int arow = min(mt * 16 + col, NEN - 1);
                    const BT aj = Aj[sj.trans ? (kn * RS + arow) : (arow * RS + kn)];
                    const BT a4 = kg == 0 ? A4[s4.trans ? (kn * RS + arow) : (arow * RS + kn)] : (BT)0;
                    acc[cc][mt] = MT_::mma(aj, (BT)bj, acc[cc][mt]);
                    acc[cc][mt] = MT_::mma(a4, (BT)b4, acc[cc][mt]);
                }
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    const int row = MT * 16 + r;
                    const OS aj = (OS)Aj[sj.trans ? (kn * RS + row) : (row * RS + kn)];
                    const OS a4 = kg == 0 ? (OS)A4[s4.trans ? (kn * RS + row) : (row * RS + kn)] : (OS)0;
                    rem[cc][r] += aj * bj + a4 * b4;
                }
            }
        }
        // rows of this lane: add into y (first touch writes)
#pragma unroll
        for (int cc = 0; cc < NC; ++cc)
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                const int c = (CSEL < 0) ? cc : CSEL;
                XS old[4];
                int dofs[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int row = mt * 16 + MT_::row(kg, i);
                    dofs[i] = (row < NEN && live) ? sLM[row * 3 + c] : -1;
                    if (PRE) {
                        old[i] = yv[PRE ? cc : 0][PRE ? mt : 0][i];
                    } else {
                        const bool add = dofs[i] >= 0 && !((fmask >> row) & 1ull);
                        old[i] = add ? reinterpret_cast<const XS *>(y + (size_t)dofs[i] * npair + pair)[ri]
                                     : (XS)0;
                    }
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const XS v = (XS)(REFINE ? accd[REFINE ? cc : 0][REFINE ? mt : 0][i] : (double)acc[cc][mt][i]);
                    if (dofs[i] >= 0) reinterpret_cast<XS *>(y + (size_t)dofs[i] * npair + pair)[ri] = old[i] + v;
                }
            }
        // VALU rows: sum the four k-lanes of each column (lanes col + 16 kg),
        // then lane kg = r stores row 16*MT + r
#pragma unroll
        for (int cc = 0; cc < NC; ++cc)
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                OS v = rem[cc][r];
                v += __shfl_xor(v, 16);
                v += __shfl_xor(v, 32);
                const int c = (CSEL < 0) ? cc : CSEL;
                const int row = MT * 16 + r;
                if (kg == (r & 3) && live) {
                    const int dof = sLM[row * 3 + c];
                    if (dof >= 0) {
                        XS *py = reinterpret_cast<XS *>(y + (size_t)dof * npair + pair) + ri;
                        const XS o = ((fmask >> row) & 1ull) ? (XS)0 : *py;
                        *py = o + (XS)v;
                    }
                }
            }
    }
}

// ---------------------------------------------------------------------------
// Batched element-block product in FP64 on the vector ALUs.
//
// On gfx950 an FP64 FMA on the vector ALUs runs at the FP64 MFMA rate (16
// per cycle per SIMD), so the MFMA form above gains nothing in FP64 and pays
// for its tile padding (NEN rows on 16-row tiles, NEN k values on 4-wide
// steps: 1.43x the useful flops at p4).  Here one wave owns RT output rows
// of one component c of one element for 64 (omega, ky) pairs, one pair per
// lane (real and imaginary part in the lane), and runs only the useful
// FMAs: 2 per (sub-block, row, node).
//
// The coefficients are the same for every lane.  They reach the FMA as its
// DPP operand: each row of 16 lanes holds the same 32 coefficients (lane
// l%16 holds numbers 2(l%16) and 2(l%16)+1 of a group of 32), and
// v_fmac_f64 with row_newbcast:j reads the value of lane j of its row, so
// one vector load feeds 64 FMAs with no broadcast instruction.  Coefficient
// layout (k_vcoef), per element e, component c, row tile t, node b: CW
// values, number u*RT + r = coefficient of sub-block kSub[c][u] between
// output row t*RT + r and input node b (zero past the last row and past
// 5 RT).  The number of row tiles NRT is a tuning choice per order (more
// tiles: fewer registers and more waves, but x read once per tile).
template <int NL, int NRT_ = (NL * NL + 24) / 25>
struct VShape {
    static constexpr int NEN = NL * NL;
    static constexpr int NRT = NRT_;                            // row tiles
    static constexpr int RT = (NEN + NRT - 1) / NRT;            // rows per tile
    static constexpr int CW = (5 * RT + 31) / 32 * 32;          // values per node
    static constexpr int NQ = CW / 32;                          // loads per node
    static constexpr int MINW = RT <= 13 ? 4 : 3;               // waves per SIMD the registers allow
    // a task (element, group of 64 pairs) is 3 NRT independent waves; a
    // workgroup is four waves of consecutive tasks (a workgroup holds
    // ceil(waves / 4) wave slots on each of the CU's four SIMDs: four-wave
    // workgroups fill every slot the registers allow)
    static constexpr int WPT = 3 * NRT;
    static constexpr int NT = 256;
};

// Sub-blocks of output component c (indices into kSub[c]) that act on input
// component c, in the order kind 0, 3 (ky^2), 4 (-omega^2): their
// coefficients are combined per pair into g = A + ky^2 B - omega^2 C before
// one complex FMA (5 instructions instead of 6); and the two others.
__device__ constexpr int kDiagU[3][3] = {{0, 3, 4}, {0, 3, 4}, {1, 3, 4}};
__device__ constexpr int kOffU[3][2] = {{1, 2}, {1, 2}, {0, 2}};

__global__ void k_vcoef(int nen, int nrt, int rt, int cw, int nelem, const double *__restrict__ blk,
                        double *__restrict__ vcf) {
    const int np = nen * nen;
    const size_t n = (size_t)nelem * 3 * nrt * nen * cw;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(i % cw);
        size_t r = i / cw;
        const int b = (int)(r % nen);
        r /= nen;
        const int t = (int)(r % nrt);
        r /= nrt;
        const int c = (int)(r % 3);
        const size_t e = r / 3;
        int u = 5, row = 0;
        if (k < 3 * rt) {
            u = kDiagU[c][k % 3];
            row = k / 3;
        } else if (k < 5 * rt) {
            u = kOffU[c][(k - 3 * rt) / rt];
            row = (k - 3 * rt) % rt;
        }
        const int a = t * rt + row;
        double v = 0.0;
        if (u < 5 && a < nen) {
            const SubBlk sb = kSub[c][u];
            const double *B = blk + (e * NBLK + sb.blk) * np;
            v = sb.trans ? B[b * nen + a] : B[a * nen + b];
        }
        vcf[i] = v;
    }
}

// acc += (value of lane J of this row of 16 lanes in cv) * op
template <int J>
__device__ __forceinline__ void fmac_bcast(double &acc, double cv, double op) {
    asm("v_fmac_f64_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
        : "+v"(acc)
        : "v"(cv), "v"(op), "i"(J));
}

// g = value of lane J of this row of 16 lanes in cv
template <int J>
__device__ __forceinline__ void mov_bcast(double &g, double cv) {
    asm("v_mov_b64_dpp %0, %1 row_newbcast:%2 row_mask:0xf bank_mask:0xf" : "=v"(g) : "v"(cv), "i"(J));
}

// Per node, coefficient k of a row tile: k < 3 RT: row k / 3, sub-block
// kDiagU[.][k % 3] (g is formed over the three, then acc += g x_c);
// 3 RT <= k < 5 RT: sub-block kOffU[.][(k - 3 RT) / RT], row (k - 3 RT) % RT.
// All instructions of one coefficient load: values 32 q + 2 j + h, h = 0, 1.
struct VOps {
    double2 xd;      // x of the component's own input component
    double2 op[2];   // operands of the two other sub-blocks
    double f3, f4;   // ky^2, -omega^2
};

template <int RT, int Q, int J>
__device__ __forceinline__ void vstep_pair(double2 *acc, double &g, const double2 &cq, const VOps &o) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        constexpr int K0 = 32 * Q + 2 * J;
        const int k = K0 + h;
        const double cv = h ? cq.y : cq.x;
        if (k < 3 * RT) {
            const int r = k / 3, s = k % 3;
            if (s == 0) {
                mov_bcast<J>(g, cv);
            } else if (s == 1) {
                fmac_bcast<J>(g, cv, o.f3);
            } else {
                fmac_bcast<J>(g, cv, o.f4);
                acc[r].x = fma(g, o.xd.x, acc[r].x);
                acc[r].y = fma(g, o.xd.y, acc[r].y);
            }
        } else if (k < 5 * RT) {
            const int u = (k - 3 * RT) / RT, r = (k - 3 * RT) % RT;
            fmac_bcast<J>(acc[r].x, cv, o.op[u].x);
            fmac_bcast<J>(acc[r].y, cv, o.op[u].y);
        }
    }
}

template <int RT, int Q, int... J>
__device__ __forceinline__ void vstep_load(double2 *acc, double &g, const double2 &cq, const VOps &o,
                                           std::integer_sequence<int, J...>) {
    (vstep_pair<RT, Q, J>(acc, g, cq, o), ...);
}

// zpad: npair zero values (read in place of the row of a fixed DOF or of a
// first touch), sink: npair values that absorb the stores of fixed DOFs, so
// that every load and store of a tile is a uniform row base plus the lane's
// pair offset and none is predicated per row.
template <int NL, int NRT, int C>
__device__ __forceinline__ void apply_v_tile(const double *__restrict__ cf, const int *__restrict__ lme,
                                             const int *__restrict__ lmo, int a0,
                                             unsigned long long fmask, int npair, int pair, bool live,
                                             double om, double kv, const double2 *__restrict__ x,
                                             double2 *__restrict__ y, const double2 *__restrict__ zpad,
                                             double2 *__restrict__ sink) {
    typedef VShape<NL, NRT> V;
    constexpr int NEN = V::NEN, RT = V::RT, CW = V::CW, NQ = V::NQ;
    const int l16 = (int)(threadIdx.x & 15);
    const double f3 = kv * kv, f4 = -om * om;
    // a lane past npair computes pair 0 and stores nothing
    const int pcl = live ? pair : 0;
    // rows of this tile: start from y unless this element touches the node
    // first (then y is written, not added to)
    double2 acc[RT];
#pragma unroll
    for (int r = 0; r < RT; ++r) {
        const int a = a0 + r;
        const int dof = a < NEN ? lmo[a * 3 + C] : -1;
        const bool add = dof >= 0 && !((fmask >> a) & 1ull);
        const double2 *row = add ? y + (size_t)dof * npair : zpad;
        acc[r] = row[pcl];
    }
    // the y values are complete before the node loop (its waits then only
    // concern the loop's own loads)
#pragma unroll
    for (int r = 0; r < RT; ++r) asm volatile("" : "+v"(acc[r].x), "+v"(acc[r].y));
    const double2 *cp = reinterpret_cast<const double2 *>(cf) + l16;
    // step b runs node b from xc, cc (loaded one step earlier) and loads node
    // b + 1 into xn, cn: the x values once the operands of node b are formed,
    // each group of 32 coefficients once its FMAs are issued, so that the
    // loaded values can take the registers just freed.  The loop runs two
    // steps per trip with the two register sets swapped (a one-set pipeline
    // needs register copies at the back edge, which wait for the loads just
    // issued).  Steps start at b = -1 (loads only): the compiler merges a
    // load before the loop with the one in it into a single load at the top
    // of the loop.  Loads past the last node read node 0 (in bounds, unused).
    auto step = [&](int b, const double2 (&cc)[NQ], double2 (&cn)[NQ], const double2 (&xc)[3],
                    double2 (&xn)[3]) {
        const int bn = b + 1 < NEN ? b + 1 : 0;
        const bool run = b >= 0 && b < NEN;
        // operands for node b
        VOps o;
        o.xd = xc[C];
        o.f3 = f3;
        o.f4 = f4;
#pragma unroll
        for (int v = 0; v < 2; ++v) {
            const SubBlk sb = kSub[C][kOffU[C][v]];
            const double2 xd = xc[sb.xcomp];
            if (sb.kind == 1) o.op[v] = make_double2(-kv * xd.y, kv * xd.x);        // +i ky x
            else if (sb.kind == 2) o.op[v] = make_double2(kv * xd.y, -kv * xd.x);   // -i ky x
            else o.op[v] = xd;
        }
        double g = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int dof = lme[bn * 3 + d];
            const double2 *row = dof >= 0 ? x + (size_t)dof * npair : zpad;
            xn[d] = row[pcl];
        }
        const double2 *cpn = cp + (size_t)bn * (CW / 2);
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            if (run) {
                if (q == 0) vstep_load<RT, 0>(acc, g, cc[q], o, std::make_integer_sequence<int, 16>());
                if (q == 1) vstep_load<RT, 1>(acc, g, cc[q], o, std::make_integer_sequence<int, 16>());
                if (q == 2) vstep_load<RT, 2>(acc, g, cc[q], o, std::make_integer_sequence<int, 16>());
                if (q == 3) vstep_load<RT, 3>(acc, g, cc[q], o, std::make_integer_sequence<int, 16>());
                if (q == 4) vstep_load<RT, 4>(acc, g, cc[q], o, std::make_integer_sequence<int, 16>());
            }
            cn[q] = cpn[q * 16];
        }
    };
    double2 xa[3], xb[3], ca[NQ], cb[NQ];
#pragma unroll
    for (int d = 0; d < 3; ++d) xa[d] = make_double2(0.0, 0.0);
#pragma unroll
    for (int q = 0; q < NQ; ++q) ca[q] = make_double2(0.0, 0.0);
#pragma unroll 1
    for (int b = -1; b < NEN; b += 2) {
        step(b, ca, cb, xa, xb);
        step(b + 1, cb, ca, xb, xa);
    }
    // read the row DOFs again rather than holding them through the loop
    asm volatile("" ::: "memory");
    if (live) {
#pragma unroll
        for (int r = 0; r < RT; ++r) {
            const int a = a0 + r;
            const int dof = a < NEN ? lmo[a * 3 + C] : -1;
            double2 *row = dof >= 0 ? y + (size_t)dof * npair : sink;
            row[pair] = acc[r];
        }
    }
}

// A task = one element of the colour and one group of 64 pairs: a wave per
// (component, row tile), 3 NRT waves that share nothing but the x rows
// they read at about the same time (one fetch from HBM serves them).  The
// pair groups of one element read the same coefficients: all tasks of an
// element go to one XCD (workgroups go to the 8 XCDs round robin).  A
// workgroup is four consecutive waves of its XCD's task list, whatever
// tasks they belong to: a workgroup holds ceil(waves / 4) wave slots on
// each of the CU's four SIMDs, so four-wave workgroups fill every slot the
// registers allow (twelve waves per CU at 164 registers), where one
// six-wave task per workgroup (p6, two row tiles) left one workgroup, six
// waves, per CU: 400x200 p6, 192 pairs, 12.1 -> 10.5 ms per launch.
// (Measured at 200x100 elements, p4, 128 pairs: 13-row tiles at four
// waves per SIMD ran 0.72 ms against 0.63.)  lm maps the element's nodes to the DOFs the input is
// read from, lm_out to the DOFs its rows go to (-1: read zero / store to
// the sink): the same array for the operator, DOF-class masks for the
// Bielak effective force.
template <int NL, int NRT>
__global__ __launch_bounds__((VShape<NL, NRT>::NT)) __attribute__((amdgpu_waves_per_eu(VShape<NL, NRT>::MINW))) void k_apply_v(
        const int *__restrict__ order, int nel, const double *__restrict__ vcf, const int *__restrict__ lm,
        const int *__restrict__ lm_out, const unsigned long long *__restrict__ first, int npair,
        const double *__restrict__ omega,
        const double *__restrict__ ky, const double2 *__restrict__ x, double2 *__restrict__ y,
        const double2 *__restrict__ zpad, double2 *__restrict__ sink) {
    typedef VShape<NL, NRT> V;
    static_assert(V::NQ <= 5, "coefficient loads per node");
    constexpr int NEN = V::NEN;
    constexpr size_t TS = (size_t)NEN * V::CW;   // values per (element, component, row tile)
    const int ngrp = (npair + 63) >> 6;
    // workgroup j runs on XCD j % 8 and takes that XCD's waves
    // 4 (j / 8) .. + 3: wave wx is wave wx % WPT of task tau = wx / WPT,
    // tau = (element block) ngrp + group, element 8 block + j % 8: all pair
    // groups of an element on one XCD
    const int j = blockIdx.x;
    const int wx = (j / 8) * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));   // wave of this XCD
    const int tau = wx / V::WPT, ct = wx % V::WPT;
    const int ei = (tau / ngrp) * 8 + j % 8, g = tau % ngrp;
    if (ei >= nel) return;
    const int e = order[ei];
    const int lane = threadIdx.x & 63;
    const unsigned long long fmask = first[e];
    const int *lme = lm + (size_t)e * 3 * NEN, *lmo = lm_out + (size_t)e * 3 * NEN;
    const int c = ct / NRT, t = ct % NRT;
    const int pair = g * 64 + lane;
    const bool live = pair < npair;
    const double om = omega[live ? pair : 0], kv = ky[live ? pair : 0];
    const double *cf = vcf + ((size_t)e * 3 * NRT + ct) * TS;
    const int a0 = t * V::RT;
    if (c == 0) apply_v_tile<NL, NRT, 0>(cf, lme, lmo, a0, fmask, npair, pair, live, om, kv, x, y, zpad, sink);
    else if (c == 1) apply_v_tile<NL, NRT, 1>(cf, lme, lmo, a0, fmask, npair, pair, live, om, kv, x, y, zpad, sink);
    else apply_v_tile<NL, NRT, 2>(cf, lme, lmo, a0, fmask, npair, pair, live, om, kv, x, y, zpad, sink);
}

// Residual of the batched product: r = f - r (r holds S u on entry) for
// [dof][pair] vectors, and per block y of dofs the partial sums of |r|^2 and
// |f|^2 for each pair: part[(y * npair + pair) * 2 + {0, 1}].  Grid
// (pair groups of 64, gy); a block's four waves take DOFs y*4 + w, stepping
// by 4 gy; lane = pair.
__global__ __launch_bounds__(256) void k_resid(int64_t ndof, int npair, const double2 *__restrict__ f,
                                               double2 *__restrict__ r, double *__restrict__ part) {
    __shared__ double red[4][64][2];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int pair = blockIdx.x * 64 + lane;
    double sr = 0.0, sf = 0.0;
    if (pair < npair) {
        for (int64_t d = (int64_t)blockIdx.y * 4 + w; d < ndof; d += (int64_t)gridDim.y * 4) {
            const size_t i = (size_t)d * npair + pair;
            const double2 fv = f[i], sv = r[i];
            const double2 rv = make_double2(fv.x - sv.x, fv.y - sv.y);
            r[i] = rv;
            sr += rv.x * rv.x + rv.y * rv.y;
            sf += fv.x * fv.x + fv.y * fv.y;
        }
    }
    red[w][lane][0] = sr;
    red[w][lane][1] = sf;
    __syncthreads();
    if (w == 0 && pair < npair) {
        const size_t o = ((size_t)blockIdx.y * npair + pair) * 2;
        part[o] = red[0][lane][0] + red[1][lane][0] + red[2][lane][0] + red[3][lane][0];
        part[o + 1] = red[0][lane][1] + red[1][lane][1] + red[2][lane][1] + red[3][lane][1];
    }
}

// out[pair * 2 + k] = sum over y = 0..gy-1 of part[(y * npair + pair) * 2 + k], in order
__global__ void k_resid_sum(int gy, int npair, const double *__restrict__ part, double *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * npair) return;
    double s = 0.0;
    for (int y = 0; y < gy; ++y) s += part[(size_t)y * 2 * npair + i];
    out[i] = s;
}

// ---------------------------------------------------------------------------
// CSR basis: V0/V1/V2 per slot (see combine), scattered from element blocks.
__device__ inline int64_t find_col(const int *__restrict__ jc, int64_t b, int64_t e, int col) {
    while (b < e) {
        int64_t m = (b + e) >> 1;
        int v = jc[m];
        if (v == col) return m;
        if (v < col) b = m + 1; else e = m;
    }
    return -1;
}

// CSR value kinds, by the components of row and column: 0 same component
// (V0 + ky^2 V1 - omega^2 V2), 1 the 1-3 coupling (V0), 2 the 1-2 and 3-2
// couplings (-i ky V0), 3 the 2-1 and 2-3 couplings (+i ky V0).  V1 and V2
// are stored only for kind 0, compactly in entry order: the kind-0 entries
// before chunk q of 64 entries number cbase[q].  One wave per row.
__global__ void k_csr_kinds(int ndof, const int64_t *__restrict__ irptr, const int *__restrict__ jcptr,
                            const signed char *__restrict__ comp, signed char *__restrict__ kind) {
    const int r = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    if (r >= ndof) return;
    const int cr = comp[r];
    for (int64_t s = irptr[r] + (threadIdx.x & 63); s < irptr[r + 1]; s += 64) {
        const int cc = comp[jcptr[s]];
        int k;
        if (cr == cc) k = 0;
        else if ((cr == 0 && cc == 2) || (cr == 2 && cc == 0)) k = 1;
        else k = (cr == 1) ? 3 : 2;
        kind[s] = (signed char)k;
    }
}

// kind-0 entries per chunk of 64 (one wave per chunk)
__global__ void k_csr_chunks(int64_t nz, const signed char *__restrict__ kind, int *__restrict__ cnt) {
    const int64_t q = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    const int64_t s = q * 64 + (threadIdx.x & 63);
    if (q * 64 >= nz) return;
    const unsigned long long msk = __ballot(s < nz && kind[s] == 0);
    if ((threadIdx.x & 63) == 0) cnt[q] = __popcll(msk);
}

// compact position of every kind-0 entry (-1 for the others)
__global__ void k_csr_cidx(int64_t nz, const int64_t *__restrict__ cbase, const signed char *__restrict__ kind,
                           int64_t *__restrict__ cidx) {
    const int64_t q = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    const int lane = threadIdx.x & 63;
    const int64_t s = q * 64 + lane;
    if (q * 64 >= nz) return;
    const bool same = s < nz && kind[s] == 0;
    const unsigned long long msk = __ballot(same);
    if (s < nz) cidx[s] = same ? cbase[q] + __popcll(msk & ((1ull << lane) - 1)) : -1;
}

__global__ void k_csr_basis(int nen, int ncol_elem, const int *__restrict__ order,
                            const double *__restrict__ blk, const int *__restrict__ lm,
                            const int64_t *__restrict__ irptr, const int *__restrict__ jcptr,
                            const int64_t *__restrict__ cidx, const char *__restrict__ cdom,
                            double *__restrict__ v0, double *__restrict__ v1, double *__restrict__ v2,
                            int *__restrict__ err) {
    const int e = order[blockIdx.x];
    if (cdom[e] == 'A') return;
    const int np = nen * nen;
    const double *B = blk + (size_t)e * NBLK * np;
    const int *L = lm + (size_t)e * 3 * nen;
    for (int i = threadIdx.x; i < 9 * np; i += blockDim.x) {
        int cp = i / np, r = i % np, a = r / nen, b = r % nen;
        int ci = cp / 3, cj = cp % 3;
        int rd = L[a * 3 + ci], cd = L[b * 3 + cj];
        if (rd < 0 || cd < 0) continue;
        int64_t s = find_col(jcptr, irptr[rd], irptr[rd + 1], cd);
        if (s < 0) { *err = 3; continue; }
        const int ab = a * nen + b, ba = b * nen + a;
        if (ci == cj) {
            int kb = ci == 0 ? B_K11 : (ci == 1 ? B_K22 : B_K33);
            int mb = ci == 1 ? B_ML2M : B_MMU;
            const int64_t j = cidx[s];
            v0[s] += B[kb * np + ab];
            v1[j] += B[mb * np + ab];
            v2[j] += B[B_MRHO * np + ab];
        } else if (ci == 0 && cj == 2) v0[s] += B[B_K13 * np + ab];
        else if (ci == 2 && cj == 0) v0[s] += B[B_K13 * np + ba];
        else if (ci == 0 && cj == 1) v0[s] += B[B_T12 * np + ab];
        else if (ci == 1 && cj == 0) v0[s] += B[B_T12 * np + ba];
        else if (ci == 1 && cj == 2) v0[s] += B[B_T23 * np + ab];
        else v0[s] += B[B_T23 * np + ba];
    }
}

// S(omega, ky) from the basis: same component: V0 + ky^2 V1 - omega^2 V2,
// 1-3 coupling: V0, 1-2/2-3 couplings: -/+ i ky V0 (block signs above).
// A stream over the entries, one wave per 128 entries (two adjacent ones
// per lane, 16-byte loads and stores; grid-stride): the compact V1/V2
// position of a kind-0 entry is the base of its 64-entry chunk (cbase) plus
// the kind-0 entries before it in the chunk (ballot and popcount).  Bytes
// per entry: kind 1 + V0 8 + the value; V1, V2 only for kind 0.
template <typename CT>
__device__ inline void csr_value(int k, double a, double k2, double w2, double kv, const double *__restrict__ v1,
                                 const double *__restrict__ v2, int64_t j, CT &o) {
    double re = 0.0, im = 0.0;
    if (k == 0) re = a + k2 * v1[j] - w2 * v2[j];
    else if (k == 1) re = a;
    else im = (k == 2 ? -kv : kv) * a;
    o.x = re;
    o.y = im;
}

template <typename CT>
__global__ void k_csr_combine(int64_t nz, const int64_t *__restrict__ cbase, const signed char *__restrict__ kind,
                              const double *__restrict__ v0, const double *__restrict__ v1,
                              const double *__restrict__ v2, double om, double kv, CT *__restrict__ out) {
    const int lane = threadIdx.x & 63;
    const double k2 = kv * kv, w2 = om * om;
    const unsigned long long lt = (1ull << lane) - 1;
    const int64_t nq = (nz + 127) / 128, step = (int64_t)gridDim.x * (blockDim.x / 64);
    for (int64_t q = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; q < nq; q += step) {
        const int64_t s = q * 128 + 2 * lane;
        const bool in0 = s < nz, in1 = s + 1 < nz;
        int ka = -1, kb = -1;
        if (in1) {
            const short kk = *reinterpret_cast<const short *>(kind + s);   // s even: 2-byte aligned
            ka = (signed char)(kk & 0xff);
            kb = (signed char)(kk >> 8);
        } else if (in0) {
            ka = kind[s];
        }
        // kind-0 entries of the first 64-entry chunk (lanes 0..31) count into
        // the second one's base; every lane's prefix starts from chunk 2q's base
        const unsigned long long ma = __ballot(ka == 0), mb = __ballot(kb == 0);
        const int64_t j = cbase[2 * q] + __popcll(ma & lt) + __popcll(mb & lt);
        if (!in0) continue;
        if (in1) {
            const double2 a = *reinterpret_cast<const double2 *>(v0 + s);
            CT o0, o1;
            csr_value(ka, a.x, k2, w2, kv, v1, v2, j, o0);
            csr_value(kb, a.y, k2, w2, kv, v1, v2, j + (ka == 0), o1);
            out[s] = o0;
            out[s + 1] = o1;
        } else {
            CT o0;
            csr_value(ka, v0[s], k2, w2, kv, v1, v2, j, o0);
            out[s] = o0;
        }
    }
}

// ---------------------------------------------------------------------------
// PML elements (CJACCQ + CM25ISOQ, cjac.f:399-581, estiff25.f:168-334),
// integrated per call and added into the assembled operator.
struct PmlPar { double omega, kv, xwidth, zwidth, rcoeff, dmax, kappa, fcent, freq, freq0; };

// haskattn.f:602-641 VCMPLX: V (1 + ln(f/f0)/(pi Q)) - i V/(2Q), real factor
// clamped at 0.1; real V when FREQ0 = 0, Q = 9999 or Q = 0
__device__ inline double2 vcmplx(double freq, double freq0, double v, double q) {
    if (freq0 == 0.0 || fabs(q - 9999.0) < 1e-5 || q == 0.0) return make_double2(v, 0.0);
    double fact = 1.0 + 1.0 / (M_PI * q) * log(freq / freq0);
    if (fact < 0.1) fact = 0.1;
    return make_double2(v * fact, -v / (2.0 * q));
}

__device__ inline double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 cdiv1(double2 z) { double d = z.x * z.x + z.y * z.y; return make_double2(z.x / d, -z.y / d); }

// (the parameters are taken by value: a reference to the kernel's argument
// struct would copy it to scratch memory)
__device__ inline double2 zdampfn(double rn, double vp, double delta, double pdmax, double rcoeff, double fcent,
                                  double omega, double kappa) {
    double dmax = (pdmax > 0.0) ? pdmax : -3.0 * vp * log(rcoeff) / (2.0 * delta);
    double d = dmax * (rn / delta) * (rn / delta);
    double alpha = M_PI * fcent * (1.0 - rn / delta);
    // kappa + |d| / (alpha + i omega), then its reciprocal (cjac.f:849-851)
    double2 t = cdiv1(make_double2(alpha, omega));
    return cdiv1(make_double2(kappa + fabs(d) * t.x, fabs(d) * t.y));
}

// Element matrices integrated per call and added into CSR order:
//  * PML elements ('A', CM25ISOQ estiff25.f:168-334 with CJACCQ's stretch)
//  * with quality factors (qp != null) every element in the density-
//    normalised dispersive form: KMDAMP (estiff25.f:511-752) and, for 'A'
//    elements, CMDAMP (estiff25.f:756-961), velocities from VCMPLX.
// Both are the KM25ISOQ bilinear forms with complex coefficients at each
// quadrature point: lambda -> L, mu -> M, rho -> R, derivatives stretched by
// 1/e_x, 1/e_z (ZDAMPFN) inside PML elements:
//   standard:   L = w lambda, M = w mu, R = w rho
//   dispersive: L = w (Vp~^2 - 2 Vs~^2), M = w Vs~^2, R = w.
// One workgroup per element of one colour.
// Stages QC quadrature points q0 .. q0+QC-1 of element e in LDS: per point
// sq[qq*9 + ..] = L (re, im), M (re, im), R, e_x (re, im), e_z (re, im) and
// per node sN, sX, sZ = N, dN/dx, dN/dz (zero past the last point).
__device__ inline void pml_stage(int nen, int nq1, int q0, int QC, int e, const double *__restrict__ shl,
                                 const double *__restrict__ wq, const double *__restrict__ xq,
                                 const int *__restrict__ ieng, const double *__restrict__ xl,
                                 const double *__restrict__ zl, const double *__restrict__ xd,
                                 const double *__restrict__ zd, bool pml, const double *__restrict__ rho,
                                 const double *__restrict__ lam, const double *__restrict__ mu,
                                 const double *__restrict__ qp, const double *__restrict__ qs, const PmlPar &par,
                                 double *sq, double *sN, double *sX, double *sZ, int *__restrict__ err) {
    const int QS = 9, nq = nq1 * nq1;
    const int *ie = ieng + 4 * e;
    for (int i = threadIdx.x; i < QC * nen; i += blockDim.x) {
        int qq = i / nen, aa = i % nen, q = q0 + qq;
        if (q >= nq) { sN[i] = sX[i] = sZ[i] = 0.0; if (aa == 0) for (int j = 0; j < QS; ++j) sq[qq * QS + j] = 0.0; continue; }
        int ix = q % nq1, iz = q / nq1;
        Geo g;
        int bad = geo_at(xq[ix], xq[iz], ie, xl, zl, g);
        if (bad && aa == 0) *err = 2;
        double rj = bad ? 1.0 : g.det;
        const double *sh = shl + ((size_t)q * nen + aa) * 3;
        sN[i] = sh[2];
        sX[i] = (sh[0] * g.dzdeta - sh[1] * g.dzdxi) / rj;
        sZ[i] = -(sh[0] * g.dxdeta - sh[1] * g.dxdxi) / rj;
        if (aa == 0) {
            double w = wq[ix] * wq[iz] * g.det, r = 0, l = 0, m = 0, xdep = 0, zdep = 0, vel = 0;
            double vsv = 0, qpv = 0, qsv = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                int gi = ie[k] - 1;
                r += g.sf[k] * rho[gi];
                l += g.sf[k] * lam[gi];
                m += g.sf[k] * mu[gi];
                if (pml) {
                    xdep += g.sf[k] * xd[gi];
                    zdep += g.sf[k] * zd[gi];
                }
                vel += g.sf[k] * sqrt((lam[gi] + 2.0 * mu[gi]) / rho[gi]);
                if (qp) {
                    vsv += g.sf[k] * sqrt(mu[gi] / rho[gi]);
                    qpv += g.sf[k] * qp[gi];
                    qsv += g.sf[k] * qs[gi];
                }
            }
            const double2 one = make_double2(1.0, 0.0);
            double2 ex = xdep > 2.22e-16
                             ? zdampfn(xdep, vel, par.xwidth, par.dmax, par.rcoeff, par.fcent, par.omega, par.kappa)
                             : one;
            double2 ez = zdep > 2.22e-16
                             ? zdampfn(zdep, vel, par.zwidth, par.dmax, par.rcoeff, par.fcent, par.omega, par.kappa)
                             : one;
            double *o = sq + qq * QS;
            if (qp) {
                const double2 zp = vcmplx(par.freq, par.freq0, vel, qpv);
                const double2 zs = vcmplx(par.freq, par.freq0, vsv, qsv);
                double2 zp2 = cmul(zp, zp), zs2 = cmul(zs, zs);
                zp2.x *= w; zp2.y *= w; zs2.x *= w; zs2.y *= w;
                o[0] = zp2.x - 2.0 * zs2.x; o[1] = zp2.y - 2.0 * zs2.y;
                o[2] = zs2.x; o[3] = zs2.y; o[4] = w;
            } else {
                o[0] = w * l; o[1] = 0.0; o[2] = w * m; o[3] = 0.0; o[4] = w * r;
            }
            o[5] = ex.x; o[6] = ex.y; o[7] = ez.x; o[8] = ez.y;
        }
    }
}

// The bilinear forms of one (a, b) node pair at one staged point, split by
// the power of ky they carry: s0 (ky^0: the stiffness, 3x3 components,
// without the mass), s1 (ky^1: the i ky couplings 12, 21, 23, 32 with i
// applied; the ky factor left out), s2 (ky^2: the mass-like diagonal
// M, L + 2M, M) and nn = N_a N_b for the -omega^2 R mass.
__device__ inline void pml_point(const double *o, double na, double nb, double xa0, double xb0, double za0,
                                 double zb0, double2 *s0, double2 *s1, double2 *s2, double &rnn) {
    const double2 L = make_double2(o[0], o[1]), M = make_double2(o[2], o[3]);
    const double R = o[4];
    const double2 ex = make_double2(o[5], o[6]), ez = make_double2(o[7], o[8]);
    double2 xa = make_double2(xa0 * ex.x, xa0 * ex.y), xb = make_double2(xb0 * ex.x, xb0 * ex.y);
    double2 za = make_double2(za0 * ez.x, za0 * ez.y), zb = make_double2(zb0 * ez.x, zb0 * ez.y);
    double2 xx = cmul(xa, xb), zz = cmul(za, zb), xz = cmul(xa, zb), zx = cmul(za, xb);
    const double2 L2 = make_double2(L.x + 2.0 * M.x, L.y + 2.0 * M.y);
    double2 t;
    t = cmul(L2, xx); s0[0].x += t.x; s0[0].y += t.y;                      // 11
    t = cmul(M, zz);  s0[0].x += t.x; s0[0].y += t.y;
    t = cmul(M, make_double2(xx.x + zz.x, xx.y + zz.y)); s0[4].x += t.x; s0[4].y += t.y;   // 22
    t = cmul(L2, zz); s0[8].x += t.x; s0[8].y += t.y;                      // 33
    t = cmul(M, xx);  s0[8].x += t.x; s0[8].y += t.y;
    t = cmul(L, xz);  s0[2].x += t.x; s0[2].y += t.y;                      // 13
    t = cmul(M, zx);  s0[2].x += t.x; s0[2].y += t.y;
    t = cmul(L, zx);  s0[6].x += t.x; s0[6].y += t.y;                      // 31
    t = cmul(M, xz);  s0[6].x += t.x; s0[6].y += t.y;
    // i ky couplings: CE12 = i ky (L xa nb - M na xb), KE(1,2) -= CE12, etc.
    const double2 Lx = cmul(L, xa), Mx = cmul(M, xb), Lz = cmul(L, zb), Mz = cmul(M, za);
    const double2 Lxb = cmul(L, xb), Mxa = cmul(M, xa), Lza = cmul(L, za), Mzb = cmul(M, zb);
    double2 t12 = make_double2(Lx.x * nb - Mx.x * na, Lx.y * nb - Mx.y * na);
    double2 t21 = make_double2(Lxb.x * na - Mxa.x * nb, Lxb.y * na - Mxa.y * nb);
    double2 t23 = make_double2(Lz.x * na - Mz.x * nb, Lz.y * na - Mz.y * nb);
    double2 t32 = make_double2(Lza.x * nb - Mzb.x * na, Lza.y * nb - Mzb.y * na);
    // times i: (x, y) -> (-y, x)
    s1[0].x += t12.y;  s1[0].y -= t12.x;    // -(i t12)
    s1[1].x -= t21.y;  s1[1].y += t21.x;    // +(i t21)
    s1[2].x -= t23.y;  s1[2].y += t23.x;    // +(i t23)
    s1[3].x += t32.y;  s1[3].y -= t32.x;    // -(i t32)
    const double nn = na * nb;
    s2[0].x += M.x * nn;  s2[0].y += M.y * nn;
    s2[1].x += L2.x * nn; s2[1].y += L2.y * nn;
    s2[2].x += M.x * nn;  s2[2].y += M.y * nn;
    rnn += R * nn;
}

// Element matrices integrated per call and added into CSR order:
//  * PML elements ('A', CM25ISOQ estiff25.f:168-334 with CJACCQ's stretch)
//  * with quality factors (qp != null) every element in the density-
//    normalised dispersive form: KMDAMP (estiff25.f:511-752) and, for 'A'
//    elements, CMDAMP (estiff25.f:756-961), velocities from VCMPLX.
// Both are the KM25ISOQ bilinear forms with complex coefficients at each
// quadrature point: lambda -> L, mu -> M, rho -> R, derivatives stretched by
// 1/e_x, 1/e_z (ZDAMPFN) inside PML elements:
//   standard:   L = w lambda, M = w mu, R = w rho
//   dispersive: L = w (Vp~^2 - 2 Vs~^2), M = w Vs~^2, R = w.
// One workgroup per element of one colour.
template <typename CT>
__global__ void k_pml(int nen, int nq1, int ncol_elem, const int *__restrict__ order,
                      const double *__restrict__ shl, const double *__restrict__ wq,
                      const double *__restrict__ xq, const int *__restrict__ ieng,
                      const double *__restrict__ xl, const double *__restrict__ zl,
                      const double *__restrict__ xd, const double *__restrict__ zd,
                      const char *__restrict__ cdom,
                      const double *__restrict__ rho, const double *__restrict__ lam,
                      const double *__restrict__ mu, const double *__restrict__ qp,
                      const double *__restrict__ qs, const int *__restrict__ lm,
                      const int64_t *__restrict__ irptr, const int *__restrict__ jcptr, PmlPar par,
                      CT *__restrict__ out, int *__restrict__ err) {
    extern __shared__ double sm[];
    const int e = order[blockIdx.x];
    const int nq = nq1 * nq1;
    // LDS: per q: L(2), M(2), R, ex(2), ez(2); then N, Nx, Nz per a for a chunk
    const int QC = 8, QS = 9;
    double *sq = sm;                 // QC*QS
    double *sN = sq + QC * QS;       // QC*nen
    double *sX = sN + QC * nen;
    double *sZ = sX + QC * nen;
    const bool pml = cdom[e] == 'A';
    const int np = nen * nen;
    const double k2 = par.kv * par.kv, w2 = par.omega * par.omega;
    for (int p0 = 0; p0 < np; p0 += blockDim.x) {
        const int pr = p0 + threadIdx.x;
        const int a = pr / nen, b = pr % nen;
        double2 s0[9], s1[4], s2[3];
        double rnn = 0.0;
#pragma unroll
        for (int i = 0; i < 9; ++i) s0[i] = make_double2(0.0, 0.0);
#pragma unroll
        for (int i = 0; i < 4; ++i) s1[i] = make_double2(0.0, 0.0);
#pragma unroll
        for (int i = 0; i < 3; ++i) s2[i] = make_double2(0.0, 0.0);
        for (int q0 = 0; q0 < nq; q0 += QC) {
            __syncthreads();
            pml_stage(nen, nq1, q0, QC, e, shl, wq, xq, ieng, xl, zl, xd, zd, pml, rho, lam, mu, qp, qs, par,
                      sq, sN, sX, sZ, err);
            __syncthreads();
            if (pr < np)
                for (int qq = 0; qq < QC; ++qq)
                    pml_point(sq + qq * QS, sN[qq * nen + a], sN[qq * nen + b], sX[qq * nen + a],
                              sX[qq * nen + b], sZ[qq * nen + a], sZ[qq * nen + b], s0, s1, s2, rnn);
        }
        if (pr < np) {
            double2 acc[9];
#pragma unroll
            for (int i = 0; i < 9; ++i) acc[i] = s0[i];
            // ky terms and the mass: S = s0 + ky s1 + ky^2 s2 - omega^2 R N N
            const int k1[4] = {1, 3, 5, 7};
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                acc[k1[i]].x += par.kv * s1[i].x;
                acc[k1[i]].y += par.kv * s1[i].y;
            }
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                acc[4 * i].x += -w2 * rnn + k2 * s2[i].x;
                acc[4 * i].y += k2 * s2[i].y;
            }
            const int *L = lm + (size_t)e * 3 * nen;
            // unrolled, so that acc stays in registers
#pragma unroll
            for (int ci = 0; ci < 3; ++ci)
#pragma unroll
                for (int cj = 0; cj < 3; ++cj) {
                    int rd = L[a * 3 + ci], cd = L[b * 3 + cj];
                    if (rd < 0 || cd < 0) continue;
                    int64_t s = find_col(jcptr, irptr[rd], irptr[rd + 1], cd);
                    if (s < 0) { *err = 3; continue; }
                    out[s].x += acc[ci * 3 + cj].x;
                    out[s].y += acc[ci * 3 + cj].y;
                }
        }
    }
}

// Elements of the batched product integrated per omega: the PML elements,
// and with quality factors (qp != null) every element in the dispersive
// form (KMDAMP / CMDAMP with VCMPLX moduli, as k_pml).  Per element (slot s
// of the launch's colour order) and distinct omega f of the batch, the
// element matrix by power of ky (CM25ISOQ with CJACCQ's stretch,
// estiff25.f:168-334):
//   S(omega_f, ky) = S0[s][f] + ky S1[s][f] + ky^2 S2[s][f]
// with rows and columns r = 3 a + i (node a, component i), NEE = 3 NEN.
// mats: [slot][nom][3][NEE][NEE] complex (S0 holds the -omega^2 mass).
// One workgroup per (element of one colour, omega).
__global__ void k_pml_split(int nen, int nq1, int slot0, int nom, const int *__restrict__ order,
                            const double *__restrict__ omf, const double *__restrict__ shl,
                            const double *__restrict__ wq, const double *__restrict__ xq,
                            const int *__restrict__ ieng, const double *__restrict__ xl,
                            const double *__restrict__ zl, const double *__restrict__ xd,
                            const double *__restrict__ zd, const char *__restrict__ cdom,
                            const double *__restrict__ rho, const double *__restrict__ lam,
                            const double *__restrict__ mu, const double *__restrict__ qp,
                            const double *__restrict__ qs, PmlPar par, double2 *__restrict__ mats,
                            int *__restrict__ err) {
    extern __shared__ double sm[];
    const int e = order[blockIdx.x], slot = slot0 + blockIdx.x, f = blockIdx.y;
    const int nq = nq1 * nq1, nee = 3 * nen;
    const int QC = 8, QS = 9;
    double *sq = sm;
    double *sN = sq + QC * QS;
    double *sX = sN + QC * nen;
    double *sZ = sX + QC * nen;
    par.omega = omf[f];
    par.freq = omf[f] / (2.0 * M_PI);
    const bool pml = cdom[e] == 'A';
    const double w2 = par.omega * par.omega;
    const int np = nen * nen;
    double2 *S0 = mats + ((size_t)slot * nom + f) * 3 * nee * nee;
    double2 *S1 = S0 + (size_t)nee * nee;
    double2 *S2 = S1 + (size_t)nee * nee;
    for (int p0 = 0; p0 < np; p0 += blockDim.x) {
        const int pr = p0 + threadIdx.x;
        const int a = pr / nen, b = pr % nen;
        double2 s0[9], s1[4], s2[3];
        double rnn = 0.0;
#pragma unroll
        for (int i = 0; i < 9; ++i) s0[i] = make_double2(0.0, 0.0);
#pragma unroll
        for (int i = 0; i < 4; ++i) s1[i] = make_double2(0.0, 0.0);
#pragma unroll
        for (int i = 0; i < 3; ++i) s2[i] = make_double2(0.0, 0.0);
        for (int q0 = 0; q0 < nq; q0 += QC) {
            __syncthreads();
            pml_stage(nen, nq1, q0, QC, e, shl, wq, xq, ieng, xl, zl, xd, zd, pml, rho, lam, mu, qp, qs, par,
                      sq, sN, sX, sZ, err);
            __syncthreads();
            if (pr < np)
                for (int qq = 0; qq < QC; ++qq)
                    pml_point(sq + qq * QS, sN[qq * nen + a], sN[qq * nen + b], sX[qq * nen + a],
                              sX[qq * nen + b], sZ[qq * nen + a], sZ[qq * nen + b], s0, s1, s2, rnn);
        }
        if (pr < np) {
#pragma unroll
            for (int i = 0; i < 3; ++i) s0[4 * i].x -= w2 * rnn;
            const double2 z = make_double2(0.0, 0.0);
#pragma unroll
            for (int ci = 0; ci < 3; ++ci)
#pragma unroll
                for (int cj = 0; cj < 3; ++cj) {
                    const size_t o = (size_t)(3 * a + ci) * nee + 3 * b + cj;
                    const int k = ci * 3 + cj;
                    S0[o] = s0[k];
                    S1[o] = (k == 1) ? s1[0] : (k == 3) ? s1[1] : (k == 5) ? s1[2] : (k == 7) ? s1[3] : z;
                    S2[o] = (ci == cj) ? s2[ci] : z;
                }
        }
    }
}

// y_b += S_e(omega_b, ky_b) x_b over the elements of one colour whose
// matrices k_pml_split holds, after the element-block product has written y
// (it stores zero rows for PML elements; for a dispersive medium y is
// zero-filled first).  Workgroup (element, group of up to 16 pairs of one
// omega): lane t % 16 is the pair, t / 16 picks rows r = t/16, t/16 + 16,
// ...; the element's x values of the group are staged in LDS.
// grp[3 g + 0..2] = (omega index, first slot in gpair, count).  Arithmetic
// in FP64 for every vector type.
template <typename XT>
__global__ __launch_bounds__(256) void k_apply_pml(int nen, int slot0, int nom, const int *__restrict__ order,
                                                   const int *__restrict__ lm, const int *__restrict__ grp,
                                                   const int *__restrict__ gpair, int npair,
                                                   const double *__restrict__ ky,
                                                   const double2 *__restrict__ mats, const XT *__restrict__ x,
                                                   XT *__restrict__ y) {
    extern __shared__ double2 xs[];   // [NEE][16]
    __shared__ int sl[3 * 7 * 7];
    const int nee = 3 * nen;
    const int e = order[blockIdx.x], slot = slot0 + blockIdx.x, g = blockIdx.y;
    const int f = grp[3 * g], first = grp[3 * g + 1], cnt = grp[3 * g + 2];
    const int pi = threadIdx.x & 15, rl = threadIdx.x >> 4;
    const bool live = pi < cnt;
    const int pair = live ? gpair[first + pi] : 0;
    for (int i = threadIdx.x; i < nee; i += blockDim.x) sl[i] = lm[(size_t)e * nee + i];
    __syncthreads();
    for (int c = rl; c < nee; c += 16) {
        const int d = sl[c];
        double2 v = make_double2(0.0, 0.0);
        if (live && d >= 0) {
            const XT t = x[(size_t)d * npair + pair];
            v = make_double2((double)t.x, (double)t.y);
        }
        xs[c * 16 + pi] = v;
    }
    __syncthreads();
    const double k = ky[pair], k2 = k * k;
    const double2 *S0 = mats + ((size_t)slot * nom + f) * 3 * nee * nee;
    const double2 *S1 = S0 + (size_t)nee * nee;
    const double2 *S2 = S1 + (size_t)nee * nee;
    for (int r = rl; r < nee; r += 16) {
        const int d = sl[r];
        const double2 *a0 = S0 + (size_t)r * nee, *a1 = S1 + (size_t)r * nee, *a2 = S2 + (size_t)r * nee;
        double ar = 0.0, ai = 0.0;
        for (int c = 0; c < nee; ++c) {
            const double2 m0 = a0[c], m1 = a1[c], m2 = a2[c];
            const double mr = m0.x + k * m1.x + k2 * m2.x, mi = m0.y + k * m1.y + k2 * m2.y;
            const double2 v = xs[c * 16 + pi];
            ar += mr * v.x - mi * v.y;
            ai += mr * v.y + mi * v.x;
        }
        if (live && d >= 0) {
            XT &o = y[(size_t)d * npair + pair];
            const double2 prev = make_double2((double)o.x, (double)o.y);
            o.x = prev.x + ar;
            o.y = prev.y + ai;
        }
    }
}

// ---------------------------------------------------------------------------
// y = S x (one wave per row)
template <typename CT, typename AT>
__global__ void k_spmv(int ndof, const int64_t *__restrict__ irptr, const int *__restrict__ jcptr,
                       const CT *__restrict__ s, const CT *__restrict__ x, CT *__restrict__ y) {
    const int wpb = blockDim.x / 64;
    const int r = blockIdx.x * wpb + threadIdx.x / 64;
    const int l = threadIdx.x & 63;
    if (r >= ndof) return;
    AT ar = 0, ai = 0;
    for (int64_t k = irptr[r] + l; k < irptr[r + 1]; k += 64) {
        CT a = s[k], b = x[jcptr[k]];
        ar += (AT)a.x * b.x - (AT)a.y * b.y;
        ai += (AT)a.x * b.y + (AT)a.y * b.x;
    }
    for (int o = 32; o > 0; o >>= 1) {
        ar += __shfl_xor(ar, o);
        ai += __shfl_xor(ai, o);
    }
    if (l == 0) { y[r].x = ar; y[r].y = ai; }
}

// Bielak effective force (asmble.f90:357-373 restricted to B/E couplings).
// Local row r is global DOF mydofs[r]-1 (identity when mydofs is null);
// out is indexed by global DOF, jcptr holds 0-based global columns.
__global__ void k_peff(int ndofl, const int *__restrict__ mydofs, const int64_t *__restrict__ irptr,
                       const int *__restrict__ jcptr, const char *__restrict__ cnp,
                       const float2 *__restrict__ s, const float2 *__restrict__ ue,
                       float2 *__restrict__ out) {
    const int wpb = blockDim.x / 64;
    const int r = blockIdx.x * wpb + threadIdx.x / 64;
    const int l = threadIdx.x & 63;
    if (r >= ndofl) return;
    const int gr = mydofs ? mydofs[r] - 1 : r;
    const char c1 = cnp[gr];
    float ar = 0.f, ai = 0.f;
    if (c1 == 'B' || c1 == 'E') {
        for (int64_t k = irptr[r] + l; k < irptr[r + 1]; k += 64) {
            int j = jcptr[k];
            char c2 = cnp[j];
            float sg = (c1 == 'B' && c2 == 'E') ? -1.f : ((c1 == 'E' && c2 == 'B') ? 1.f : 0.f);
            if (sg != 0.f) {
                float2 a = s[k], b = ue[j];
                ar += sg * (a.x * b.x - a.y * b.y);
                ai += sg * (a.x * b.y + a.y * b.x);
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        ar += __shfl_xor(ar, o);
        ai += __shfl_xor(ai, o);
    }
    if (l == 0) out[gr] = make_float2(ar, ai);
}

// ---------------------------------------------------------------------------
// Adjoint gradient densities.  For fields u (forward) and l (adjoint) with
// d/dy -> -i ky and test-side conjugation:
//   d_lambda = Re[conj(div l) div u],  div u = u1,x + u3,z - i ky u2
//   d_mu     = Re[conj(grad l) : (grad u + grad u^T)]
//   d_rho    = -omega^2 Re[conj(l) . u]
// accumulated over the wavefield pairs at each quadrature point, then
// projected on the element's four anchor hat functions.

// Adjoint gradient densities by sum factorisation (every order, p2..p6).
// The shape functions are tensor products N_a(q) = v_i(xi)
// v_j(eta), a = j N1 + i (dshl.f:149-175), so per trip, field w (u or l)
// and component c
//   T1[eta][i] = sum_j v_j(eta) w_(j,i,c),   T2[eta][i] = sum_j v'_j(eta) w_(j,i,c)
//   U = sum_i v_i(xi) T1,  U_xi = sum_i v'_i(xi) T1,  U_eta = sum_i v_i(xi) T2
// : N1 Q1 (2 N1 + 3 Q1) complex MACs instead of 3 N1^2 Q1^2 (p4: 4.6x
// fewer).  The second stage runs on rows of 16 lanes, one row per eta
// point and one lane per xi point; lane i of a row holds T1, T2 of node i
// and v_fmac_f64 with row_newbcast:i hands them to the whole row.  One
// workgroup per element, the geometry at q computed once per element; the
// 1-D tables are not flushed to zero as DSHL's
// products are (a difference below 2.2e-16 per shape value).  The element
// values of the next pass of triples are loaded into registers while the
// current pass computes, the forward field's gradient is reduced to its
// symmetrised form before the adjoint field's components are formed one at
// a time, and the quadrature weights are recomputed for the projection
// instead of kept in LDS: 128 registers and 37 KB of LDS per workgroup at
// p4, four workgroups per CU.  At p6 the FP64 form takes two points per
// lane and one triple per pass: three waves, 168 registers and 29.5 KB of
// LDS, four workgroups (twelve waves) per CU.
template <int NL, int I, int... J>
__device__ __forceinline__ void sf_row_sums(double2 &U, double2 &UX, double2 &UE, const double2 &T1,
                                            const double2 &T2, const double *vx, const double *dx,
                                            std::integer_sequence<int, J...>) {
    ((fmac_bcast<J>(U.x, T1.x, vx[J]), fmac_bcast<J>(U.y, T1.y, vx[J]), fmac_bcast<J>(UX.x, T1.x, dx[J]),
      fmac_bcast<J>(UX.y, T1.y, dx[J]), fmac_bcast<J>(UE.x, T2.x, vx[J]), fmac_bcast<J>(UE.y, T2.y, vx[J])),
     ...);
}

// The same sums with the lane's 1-D values read from an LDS table instead
// of registers: sv[j Q1] = v_j(xi), sv[(NL + j) Q1] = v'_j(xi) at the
// lane's xi point.
template <int NL, int Q1, int... J>
__device__ __forceinline__ void sf_row_sums_lds(double2 &U, double2 &UX, double2 &UE, const double2 &T1,
                                                const double2 &T2, const double *sv,
                                                std::integer_sequence<int, J...>) {
    ((fmac_bcast<J>(U.x, T1.x, sv[J * Q1]), fmac_bcast<J>(U.y, T1.y, sv[J * Q1]),
      fmac_bcast<J>(UX.x, T1.x, sv[(NL + J) * Q1]), fmac_bcast<J>(UX.y, T1.y, sv[(NL + J) * Q1]),
      fmac_bcast<J>(UE.x, T2.x, sv[J * Q1]), fmac_bcast<J>(UE.y, T2.y, sv[J * Q1])),
     ...);
}

// The same sums from an interleaved LDS table: si[J] = (v_J, v'_J) at the
// lane's xi point (one 16-byte read per node instead of two 8-byte ones).
template <int... J>
__device__ __forceinline__ void sf_row_sums_lds2(double2 &U, double2 &UX, double2 &UE, const double2 &T1,
                                                 const double2 &T2, const double2 *si,
                                                 std::integer_sequence<int, J...>) {
    ((fmac_bcast<J>(U.x, T1.x, si[J].x), fmac_bcast<J>(U.y, T1.y, si[J].x), fmac_bcast<J>(UX.x, T1.x, si[J].y),
      fmac_bcast<J>(UX.y, T1.y, si[J].y), fmac_bcast<J>(UE.x, T2.x, si[J].x), fmac_bcast<J>(UE.y, T2.y, si[J].x)),
     ...);
}

// R = double: FP64 fields and arithmetic.  R = float: complex64 fields, the
// interpolation and the densities of each triple in FP32, their sums over
// triples, quadrature points and corners in FP64 (the FP32 gradient); its
// second stage reads T1, T2 from LDS (no DPP broadcast).
template <typename R> struct CxOf;
template <> struct CxOf<double> { typedef double2 T; };
template <> struct CxOf<float> { typedef float2 T; };

// a[p] and a[p] += v for a register array of one or two entries and a
// run-time p, by selects (an indexed register array would be placed in
// scratch memory)
template <typename V, int N>
__device__ __forceinline__ V sel(const V (&a)[N], int p) {
    static_assert(N == 1 || N == 2, "one or two points per lane");
    if constexpr (N == 1) return a[0]; else return p ? a[1] : a[0];
}
template <int N>
__device__ __forceinline__ void add_sel(double (&a)[N], int p, double v) {
    static_assert(N == 1 || N == 2, "one or two points per lane");
    if constexpr (N == 1) {
        a[0] += v;
    } else {
        const double a0 = a[0] + v, a1 = a[1] + v;
        a[0] = p ? a[0] : a0;
        a[1] = p ? a1 : a[1];
    }
}

// Launch shape.  One point per lane: rows of 16 lanes over the eta points
// (at Q1 = 17 plus a column of 17 lanes for xi point 16), four waves up to
// p5 and five at p6.  The FP64 form at p6 takes two points per lane in its
// first two waves (row r < 8 holds eta points r and r + 9) and one in the
// third (eta 8, then the xi = 16 column), which therefore runs the point
// stage once per triple instead of twice; three waves.  Measured residency at the
// FP64 form's three waves per SIMD: one five-wave workgroup per CU (a
// workgroup of four or more waves holds ceil(waves / 4) slots on every
// SIMD), four three-wave ones.  Up to p4 FP64 keeps four workgroups per
// CU; FP32 five waves per SIMD (its registers allow it).
template <int NL, typename R>
struct GradOcc {
    static constexpr int Q1 = 2 * NL + 3;
    static constexpr bool F64 = std::is_same<R, double>::value;
    static constexpr int PPL = (F64 && Q1 > 16) ? 2 : 1;   // quadrature points per lane
    static constexpr int NT = PPL == 2 ? 192 : (Q1 > 16 ? 320 : 256);
    // waves per SIMD the registers must allow
    static constexpr int WPE = !F64 ? 5 : (NL <= 5 ? 4 : 3);
    // triples per pass
    static constexpr int T = PPL == 2 ? 1 : 2;
};

template <int NL, int T, typename R = double>
__global__ __launch_bounds__((GradOcc<NL, R>::NT)) __attribute__((amdgpu_waves_per_eu(GradOcc<NL, R>::WPE))) void k_grad_sf(
    const double *__restrict__ t1d, const double *__restrict__ wq, const double *__restrict__ xq,
    const int *__restrict__ ieng, const double *__restrict__ xl, const double *__restrict__ zl,
    const char *__restrict__ cdom, const int *__restrict__ lm, int ntrip, int64_t ndof,
    const double *__restrict__ omega, const double *__restrict__ ky, const typename CxOf<R>::T *__restrict__ u,
    const typename CxOf<R>::T *__restrict__ l, double *__restrict__ part, int *__restrict__ err) {
    typedef typename CxOf<R>::T C2;
    constexpr bool F64 = std::is_same<R, double>::value;
    constexpr int N1 = NL, Q1 = 2 * NL + 3, nen = NL * NL, nee = 3 * nen;
    // Q1 = 17 (p6): rows of 16 lanes hold xi points 0..15 of H eta points
    // each (PPL per lane) and the H lanes after them xi point 16
    constexpr bool OVF = Q1 > 16;
    constexpr int PPL = GradOcc<NL, R>::PPL, NT = GradOcc<NL, R>::NT;
    static_assert(PPL == 2 || 16 * Q1 + (OVF ? Q1 : 0) <= NT, "a lane for every point");
    constexpr int NW = T * 2 * nee;            // element values gathered per pass
    constexpr int NG = (NW + NT - 1) / NT;     // of them per thread
    static_assert(Q1 <= 17, "one row of 16 lanes per eta point, plus one column");
    extern __shared__ double sm[];
    C2 *sw = reinterpret_cast<C2 *>(sm);                               // [T][2][nee]
    C2 *sT = sw + NW;                                                  // [T][2][3][Q1][N1][2]
    R *s1 = reinterpret_cast<R *>(sT + T * 2 * 3 * Q1 * N1 * 2);       // [2][N1][Q1]
    R *sG = s1 + 2 * N1 * Q1;                                          // FP32: [Q1*Q1][4]
    // two points per lane (FP64, p6): the 1-D values also interleaved,
    // si[q N1 + j] = (v_j(q), v'_j(q))
    constexpr bool XL = PPL == 2;
    double2 *si = reinterpret_cast<double2 *>(s1 + 2 * N1 * Q1);
    __shared__ double red[NT / 64][12];
    const int e = blockIdx.x, tid = threadIdx.x;
    if (cdom[e] == 'A') {   // PML elements are outside the non-PML gradient path
        if (tid < 12) part[(size_t)e * 12 + tid] = 0.0;
        return;
    }
    const int *ie = ieng + 4 * e;
    const int *L = lm + (size_t)e * nee;
    for (int i = tid; i < 2 * N1 * Q1; i += NT) s1[i] = (R)t1d[i];   // v_j(eta_q), then v'_j(eta_q)
    if constexpr (XL)
        for (int i = tid; i < N1 * Q1; i += NT) {
            const int q = i / N1, j = i % N1;
            si[i] = make_double2(t1d[j * Q1 + q], t1d[(N1 + j) * Q1 + q]);
        }
    if constexpr (!F64) {
        // FP32: the inverse Jacobian at every point in LDS (registers are the
        // FP32 form's occupancy limit; the FP64 form keeps its point's
        // values in registers instead, its LDS being the limit)
        for (int q = tid; q < Q1 * Q1; q += NT) {
            Geo g;
            if (geo_at(xq[q % Q1], xq[q / Q1], ie, xl, zl, g)) { *err = 2; g.det = 1.0; }
            R *G = sG + 4 * q;
            G[0] = (R)(g.dzdeta / g.det); G[1] = (R)(-g.dzdxi / g.det); G[2] = (R)(-g.dxdeta / g.det);
            G[3] = (R)(g.dxdxi / g.det);
        }
    }
    // the gather slots of this thread are the same in every pass: slot r is
    // value tid + r NT of [T][2][nee]; its DOF (or -1) is fixed per element
    int sd[NG];
#pragma unroll
    for (int r = 0; r < NG; ++r) {
        const int i = tid + r * NT;
        sd[r] = (i < NW) ? L[i % nee] : -1;
    }
    // the values of the next pass are loaded into registers before this
    // pass's arithmetic and written to LDS after it
    C2 pre[NG];
    auto gather = [&](int t0) {
#pragma unroll
        for (int r = 0; r < NG; ++r) {
            const int i = tid + r * NT;
            const int tt = i / (2 * nee), f = (i / nee) % 2, t = t0 + tt;
            const C2 *w = f ? l : u;
            pre[r] = (sd[r] >= 0 && t < ntrip) ? w[(size_t)t * ndof + sd[r]] : C2{R(0), R(0)};
        }
    };
    gather(0);
    // second stage: row = eta point, lane = xi point (clamped when idle: an
    // idle lane computes a copy and contributes nothing)
    const int row = tid >> 4, x16 = tid & 15;
    int qe[PPL], qx;
    bool act[PPL], ovl;
    // points of this lane's wave: PPL, or one for the last wave of the
    // two-point form
    int npw = PPL;
    if constexpr (PPL == 1) {
        // rows 0..Q1-1: eta = row; at Q1 = 17 the lanes 16 Q1 .. 17 Q1 - 1
        // hold xi = 16 of eta 0..16 (clamped copies where idle)
        ovl = OVF && tid >= 16 * Q1 && tid < 17 * Q1;
        qe[0] = ovl ? tid - 16 * Q1 : min(row, Q1 - 1);
        qx = ovl ? 16 : min(x16, Q1 - 1);
        act[0] = (row < Q1 && x16 < Q1) || ovl;
    } else {
        // Q1 = 17, three waves: rows 0..7 (waves 0, 1) hold eta r and r + 9,
        // xi 0..15; the last wave one point per lane: row 8 eta 8, then
        // 17 lanes the xi = 16 column (eta 0..16), the rest idle
        static_assert(PPL == 2 && Q1 == 17 && NT == 192, "the two-point layout is the p6 one");
        const bool two = tid < 128;
        ovl = tid >= 144 && tid < 161;
        const int eb = two ? row : (ovl ? tid - 144 : 8);
        qe[0] = eb;
        qe[1] = two ? row + 9 : eb;
        qx = ovl ? 16 : x16;
        act[0] = two || row == 8 || ovl;
        act[1] = two;
        npw = two ? 2 : 1;
    }
    const int ni = min(x16, N1 - 1);
    // the 1-D values at the lane's xi point: in registers, or with two
    // points per lane read from the LDS table si (registers are that form's
    // occupancy limit)
    R vx[N1], dx[N1];
    if constexpr (!XL) {
#pragma unroll
        for (int i = 0; i < N1; ++i) { vx[i] = (R)t1d[i * Q1 + qx]; dx[i] = (R)t1d[(N1 + i) * Q1 + qx]; }
    }
    // FP64: the inverse Jacobian at the lane's points, kept in registers
    R rx[PPL], sx[PPL], rz[PPL], sz[PPL];
#pragma unroll
    for (int p = 0; p < PPL; ++p) {
        rx[p] = sx[p] = rz[p] = sz[p] = R(0);
        if constexpr (F64) {
            Geo g;
            if (geo_at(xq[qx], xq[qe[p]], ie, xl, zl, g)) { *err = 2; g.det = 1.0; }
            rx[p] = g.dzdeta / g.det; sx[p] = -g.dzdxi / g.det; rz[p] = -g.dxdeta / g.det; sz[p] = g.dxdxi / g.det;
        }
    }
    double Dl[PPL], Dm[PPL], Dr[PPL];
#pragma unroll
    for (int p = 0; p < PPL; ++p) Dl[p] = Dm[p] = Dr[p] = 0.0;
    for (int t0 = 0; t0 < ntrip; t0 += T) {
        // no barrier before these writes: every wave passed the previous
        // pass's second barrier after its last read of sw, and the pass's
        // first writes of sT come after the barrier below, which no wave
        // reaches before its previous second stage is done with sT
#pragma unroll
        for (int r = 0; r < NG; ++r)
            if (tid + r * NT < NW) sw[tid + r * NT] = pre[r];
        __syncthreads();   // sw (and s1) written
        // first pass, T1 and T2 for every (fc, eta point, node column i),
        // fc = (tt*2 + f)*3 + c.  FP64: a thread per (fc, i, third of the
        // eta points) keeps the column's N1 element values in registers and
        // sweeps its eta points (each value read from LDS once per third);
        // FP32: a thread per (fc, eta point, i), fewer registers (five waves
        // per SIMD).
        if constexpr (F64) {
            constexpr int QS3 = (Q1 + 2) / 3;
            for (int it = tid; it < T * 2 * 3 * N1 * 3; it += NT) {
                const int third = it % 3, i = (it / 3) % N1, fc = it / (3 * N1);
                const int c = fc % 3;
                const C2 *w = sw + (fc / 3) * nee;
                C2 a[N1];
#pragma unroll
                for (int j = 0; j < N1; ++j) a[j] = w[(j * N1 + i) * 3 + c];
                const int q0 = third * QS3, q1 = min(Q1, q0 + QS3);
                for (int qy = q0; qy < q1; ++qy) {
                    C2 t1 = {R(0), R(0)}, t2 = t1;
#pragma unroll
                    for (int j = 0; j < N1; ++j) {
                        R vv, dd;
                        if constexpr (XL) {
                            const double2 vd = si[qy * N1 + j];
                            vv = vd.x; dd = vd.y;
                        } else {
                            vv = s1[j * Q1 + qy]; dd = s1[(N1 + j) * Q1 + qy];
                        }
                        t1.x += vv * a[j].x; t1.y += vv * a[j].y;
                        t2.x += dd * a[j].x; t2.y += dd * a[j].y;
                    }
                    C2 *o = sT + (size_t)((fc * Q1 + qy) * N1 + i) * 2;
                    o[0] = t1;
                    o[1] = t2;
                }
            }
        } else {
            for (int it = tid; it < T * 2 * 3 * Q1 * N1; it += NT) {
                const int i = it % N1, qy = (it / N1) % Q1, fc = it / (N1 * Q1);
                const int c = fc % 3;
                const C2 *w = sw + (fc / 3) * nee;
                C2 t1 = {R(0), R(0)}, t2 = t1;
#pragma unroll
                for (int j = 0; j < N1; ++j) {
                    const C2 a = w[(j * N1 + i) * 3 + c];
                    const R vv = s1[j * Q1 + qy], dd = s1[(N1 + j) * Q1 + qy];
                    t1.x += vv * a.x; t1.y += vv * a.y;
                    t2.x += dd * a.x; t2.y += dd * a.y;
                }
                C2 *o = sT + (size_t)((fc * Q1 + qy) * N1 + i) * 2;
                o[0] = t1;
                o[1] = t2;
            }
        }
        R kt[T], wt[T];
#pragma unroll
        for (int tt = 0; tt < T; ++tt) {
            const int t = min(t0 + tt, ntrip - 1);
            kt[tt] = (R)ky[t];
            wt[tt] = (R)omega[t];
        }
        __syncthreads();
        if (t0 + T < ntrip) gather(t0 + T);
        if constexpr (!F64) {
            const R *G = sG + 4 * (qx + Q1 * qe[0]);
            rx[0] = G[0]; sx[0] = G[1]; rz[0] = G[2]; sz[0] = G[3];
        }
#pragma unroll
        for (int tt = 0; tt < T; ++tt) {
          if (t0 + tt >= ntrip) break;
          // the lane's points one after the other (a loop, not unrolled: two
          // points' live values at once would exceed the registers)
#pragma unroll 1
          for (int p = 0; p < npw; ++p) {
            const R k = kt[tt], w = wt[tt];
            const int qp = sel(qe, p);
            const R rxp = sel(rx, p), sxp = sel(sx, p), rzp = sel(rz, p), szp = sel(sz, p);
            // gradient of component c of field f at this point: [d/dx, d/dy, d/dz]
            auto grad = [&](int f, int c, C2 &U, C2 *Gc) {
                const C2 *o = sT + (size_t)((((tt * 2 + f) * 3 + c) * Q1 + qp) * N1 + ni) * 2;
                const C2 T1 = o[0], T2 = o[1];
                C2 UXI = {R(0), R(0)}, UET = UXI;
                U = UXI;
                // the FP64 form broadcasts T1, T2 along the row by DPP; the
                // FP32 form, and the xi = 16 column, read the row's T1, T2
                // from LDS
                bool bc = !ovl;
                if constexpr (F64) {
                    if (bc && XL)
                        sf_row_sums_lds2(U, UXI, UET, T1, T2, si + qx * N1, std::make_integer_sequence<int, N1>());
                    else if (bc)
                        sf_row_sums<NL, 0>(U, UXI, UET, T1, T2, (const double *)vx, (const double *)dx,
                                           std::make_integer_sequence<int, N1>());
                } else {
                    bc = false;
                }
                if (!bc) {
                    const C2 *r = o - 2 * ni;
#pragma unroll
                    for (int j = 0; j < N1; ++j) {
                        const C2 a1 = r[2 * j], a2 = r[2 * j + 1];
                        const R vj = XL ? si[qx * N1 + j].x : vx[j], dj = XL ? si[qx * N1 + j].y : dx[j];
                        U.x += vj * a1.x; U.y += vj * a1.y;
                        UXI.x += dj * a1.x; UXI.y += dj * a1.y;
                        UET.x += vj * a2.x; UET.y += vj * a2.y;
                    }
                }
                Gc[0] = C2{rxp * UXI.x + sxp * UET.x, rxp * UXI.y + sxp * UET.y};
                Gc[1] = C2{k * U.y, -k * U.x};   // d/dy = -i ky
                Gc[2] = C2{rzp * UXI.x + szp * UET.x, rzp * UXI.y + szp * UET.y};
            };
            // forward field: its values, divergence and symmetrised gradient
            // S[i][j] = G_ij + G_ji; then the adjoint field's components one
            // at a time against them
            C2 GU[3][3], Uv[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) grad(0, c, Uv[c], GU[c]);
            const C2 du = C2{GU[0][0].x + GU[1][1].x + GU[2][2].x, GU[0][0].y + GU[1][1].y + GU[2][2].y};
            C2 S[3][3];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) S[i][j] = C2{GU[i][j].x + GU[j][i].x, GU[i][j].y + GU[j][i].y};
            C2 dv = {R(0), R(0)};
            R sm2 = R(0), vu = R(0);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                C2 V, GV[3];
                grad(1, c, V, GV);
                dv = c ? C2{dv.x + GV[c].x, dv.y + GV[c].y} : GV[0];
#pragma unroll
                for (int j = 0; j < 3; ++j) sm2 += GV[j].x * S[c][j].x + GV[j].y * S[c][j].y;
                vu += V.x * Uv[c].x + V.y * Uv[c].y;
            }
            add_sel(Dl, p, (double)(dv.x * du.x + dv.y * du.y));
            add_sel(Dm, p, (double)sm2);
            add_sel(Dr, p, -(double)w * (double)w * (double)vu);
          }
        }
    }
    // projection on the anchor hat functions (weights recomputed here, as the
    // quadrature loop above computes them) and block reduction of the 12 sums
    double g12[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) g12[i] = 0.0;
#pragma unroll
    for (int p = 0; p < PPL; ++p) {
        double fw[4] = {0.0, 0.0, 0.0, 0.0};
        if (act[p]) {
            Geo g;
            if (geo_at(xq[qx], xq[qe[p]], ie, xl, zl, g)) g.det = 1.0;
            const double wq2 = wq[qx] * wq[qe[p]] * g.det;
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) fw[kk] = wq2 * g.sf[kk];
        }
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            g12[kk * 3 + 0] += fw[kk] * Dl[p];
            g12[kk * 3 + 1] += fw[kk] * Dm[p];
            g12[kk * 3 + 2] += fw[kk] * Dr[p];
        }
    }
#pragma unroll
    for (int i = 0; i < 12; ++i)
        for (int o = 32; o > 0; o >>= 1) g12[i] += __shfl_xor(g12[i], o);
    if ((tid & 63) == 0)
#pragma unroll
        for (int i = 0; i < 12; ++i) red[tid >> 6][i] = g12[i];
    __syncthreads();
    if (tid < 12) {
        double s = 0.0;
        for (int wv = 0; wv < NT / 64; ++wv) s += red[wv][tid];
        part[(size_t)e * 12 + tid] = s;
    }
}

// per-anchor sum of element partials and chain rule to (Vp, Vs, rho)
// Anchors of PML elements (fixa[a] != 0) hold fixed parameters: their
// entries receive nothing.
__global__ void k_grad_gather(int nnpg, int ncon, const int *__restrict__ mconn,
                              const int *__restrict__ ieng, const double *__restrict__ part,
                              const double *__restrict__ rho, const double *__restrict__ lam,
                              const double *__restrict__ mu, const char *__restrict__ fixa, int chain,
                              double *__restrict__ g) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nnpg || fixa[a]) return;
    double gl = 0, gm = 0, gr = 0;
    for (int c = 0; c < ncon; ++c) {
        int e = mconn[(size_t)c * nnpg + a];
        if (e <= 0) break;
        --e;
        for (int k = 0; k < 4; ++k)
            if (ieng[4 * e + k] - 1 == a) {
                gl += part[(size_t)e * 12 + k * 3 + 0];
                gm += part[(size_t)e * 12 + k * 3 + 1];
                gr += part[(size_t)e * 12 + k * 3 + 2];
            }
    }
    if (chain) {
        double r = rho[a], vp = sqrt((lam[a] + 2.0 * mu[a]) / r), vs = sqrt(mu[a] / r);
        g[3 * a + 0] += gl * 2.0 * r * vp;
        g[3 * a + 1] += gl * (-4.0 * r * vs) + gm * 2.0 * r * vs;
        g[3 * a + 2] += gl * (vp * vp - 2.0 * vs * vs) + gm * vs * vs + gr;
    } else {
        g[3 * a + 0] += gl;
        g[3 * a + 1] += gm;
        g[3 * a + 2] += gr;
    }
}

// ---------------------------------------------------------------------------
// Pseudo-Hessian diagonal, element part.  For corner c of element e and the
// anchor's parameters p = (Vp, Vs, rho), the element rows of the column
// dS/dm_{a,p} u are
//   r_{b,i} = sum_q w_q N_c(q) T_{b,i}(q)
// with the test contractions of the lambda, mu and rho terms
//   T^l_{b,i}   = conj(D_{b,i}) div u         D_{b,x} = Nx_b, D_{b,y} = -i ky N_b, D_{b,z} = Nz_b
//   T^m_{b,i}   = sum_k conj(d_k N_b) (G_ik + G_ki),  G_ik = d_k u_i, d_y = -i ky
//   T^rho_{b,i} = -omega^2 N_b u_i
// (the same bilinear forms as the gradient densities: lam^H r = the
// gradient density), then the chain rule at the anchor:  Vp: 2 rho Vp T^l;
// Vs: -4 rho Vs T^l + 2 rho Vs T^m;  rho: (Vp^2 - 2 Vs^2) T^l + Vs^2 T^m + T^rho.
// out[((e*4 + c)*3 + p)*nee + 3b + i].

// fmac_bcast as a volatile statement: kept in program order after the
// first pair of its row sum (fmac2_bcast_first, which carries the wait
// states)
template <int J>
__device__ __forceinline__ void fmac_bcast_ordered(double &acc, double cv, double op) {
    asm volatile("v_fmac_f64_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
                 : "+v"(acc)
                 : "v"(cv), "v"(op), "i"(J));
}

// The first two FMAs of a row sum whose operand F was just written by
// vector-ALU instructions: gfx950 needs two wait states between a VALU
// write of a VGPR and a DPP read of it, and the compiler does not look into
// inline assembly, so the s_nop is part of the same statement.
template <int J>
__device__ __forceinline__ void fmac2_bcast_first(double2 &acc, const double2 &F, double c) {
    asm volatile(
        "s_nop 1\n\t"
        "v_fmac_f64_dpp %0, %2, %4 row_newbcast:%5 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f64_dpp %1, %3, %4 row_newbcast:%5 row_mask:0xf bank_mask:0xf"
        : "+v"(acc.x), "+v"(acc.y)
        : "v"(F.x), "v"(F.y), "v"(c), "i"(J));
}

// Row sum over the xi points of a row of 16 lanes: acc += sum_J c[J] F(lane J).
template <int J0, int... J>
__device__ __forceinline__ void sf_row_dot(double2 &acc, const double2 &F, const double *c,
                                           std::integer_sequence<int, J0, J...>) {
    fmac2_bcast_first<J0>(acc, F, c[J0]);
    ((fmac_bcast_ordered<J>(acc.x, F.x, c[J]), fmac_bcast_ordered<J>(acc.y, F.y, c[J])), ...);
}

// The element part by sum factorisation (every order, p2..p6).  With
// N_b = v_bi(xi) v_bj(eta) and the anchor hat
// functions sf_c = h_cx(xi) h_cz(eta) (h_0 = (1 - s)/2, h_1 = (1 + s)/2;
// corners c = 0..3 at (cx, cz) = (0,0), (1,0), (1,1), (0,1)), every column
// entry is
//   sum_q w sf_c [ dN_b/dxi A + dN_b/deta B + N_b C ]
//   = sum_eta h_cz v_bj P + h_cz v'_bj R,
//   P[cx][eta][bi] = sum_xi h_cx (v'_bi A + v_bi C),  R = sum_xi h_cx v_bi B,
// where (A, B, C) per quadrature point are, for the density l (i = x, y, z):
// (rx div, sx div, 0), (0, 0, i ky div), (rz div, sz div, 0); for m:
// (rx S_ix + rz S_iz, sx S_ix + sz S_iz, i ky S_iy); for the rho term:
// (0, 0, -omega^2 u_i), all times the quadrature weight w (then the chain
// rule above).  One workgroup per element.  Stages: the fields and their gradients at
// the quadrature points (rows of 16 lanes, one per eta point, as in
// k_grad_sf); per channel, lane (cx, bi) of each row forms P and R from the
// row's points by DPP broadcast; then one thread per (corner, node,
// component) contracts over eta.  14 planes of P and R (9 P, 5 R) per
// (cx, eta, bi).
template <int NL>
__global__ __launch_bounds__((2 * NL + 3 > 16 ? 320 : 256), (NL <= 5 ? 4 : 3)) void k_fcols_sf(const double *__restrict__ t1d, const double *__restrict__ wq,
                                                  const double *__restrict__ xq, const int *__restrict__ ieng,
                                                  const double *__restrict__ xl, const double *__restrict__ zl,
                                                  const char *__restrict__ cdom, const int *__restrict__ lm,
                                                  const double *__restrict__ rho, const double *__restrict__ lam,
                                                  const double *__restrict__ mu, double omega, double kv,
                                                  const double2 *__restrict__ u, double2 *__restrict__ out,
                                                  int *__restrict__ err) {
    constexpr int N1 = NL, Q1 = 2 * NL + 3, nen = NL * NL, nee = 3 * nen;
    // Q1 = 17 (p6): rows 0..16 of 16 lanes hold xi points 0..15 and the 17
    // lanes after them xi point 16 of eta rows 0..16 (five waves); that
    // column's 17 point fields go through LDS to the row sums
    constexpr bool OVF = Q1 > 16;
    constexpr int NT = OVF ? 320 : 256, QD = OVF ? 16 : Q1;   // QD: xi points summed by DPP
    constexpr int REG = (nee + 6 * Q1 * N1 > 28 * Q1 * N1) ? nee + 6 * Q1 * N1 : 28 * Q1 * N1;
    static_assert(Q1 <= 17 && 2 * N1 <= 16, "one row of 16 lanes per eta point, plus one column");
    extern __shared__ double sm[];
    double *s1 = sm;                                  // [2][N1][Q1] v_j(s_q), v'_j(s_q); [2][Q1] h_0, h_1
    double *sh = s1 + 2 * N1 * Q1;
    double2 *su = reinterpret_cast<double2 *>(sh + 2 * Q1);   // nee element values
    double2 *sT = su + nee;                                   // [3][Q1][N1][2] T1, T2
    double2 *sP = su;   // [14][2][Q1][N1], written once su and sT are read
    double2 *sF = su + REG;   // [17][Q1] point fields of the xi = 16 column (Q1 = 17)
    // Q1 = 17: the derivative row-sum coefficients h_cx v'_bi in LDS
    // [2: cx][N1][Q1] (121 instead of 160 registers: two five-wave
    // workgroups per CU instead of one; below p6 they stay in registers)
    double *sc = reinterpret_cast<double *>(sF + (OVF ? 17 * Q1 : 0));
    const int e = blockIdx.x, tid = threadIdx.x;
    double2 *o = out + (size_t)e * 12 * nee;
    if (cdom[e] == 'A') {   // PML elements are outside the non-PML gradient path
        for (int i = tid; i < 12 * nee; i += NT) o[i] = make_double2(0.0, 0.0);
        return;
    }
    const int *ie = ieng + 4 * e;
    const int *L = lm + (size_t)e * nee;
    // rows of 16 lanes: row = eta point, lane = xi point (clamped copies
    // where idle: they are never read or stored); the geometry of the lane's
    // point first, so that its loads overlap the stages before its use
    const int row = tid >> 4, x16 = tid & 15;
    const bool ovl = OVF && tid >= 16 * Q1 && tid < 17 * Q1;   // a lane of the xi = 16 column
    const int qe = ovl ? tid - 16 * Q1 : min(row, Q1 - 1), qx = ovl ? 16 : min(x16, Q1 - 1);
    const int ni = min(x16, N1 - 1);
    Geo g;
    if (geo_at(xq[qx], xq[qe], ie, xl, zl, g)) { *err = 2; g.det = 1.0; }
    const double rx = g.dzdeta / g.det, sx = -g.dzdxi / g.det, rz = -g.dxdeta / g.det, sz = g.dxdxi / g.det;
    const double w = wq[qx] * wq[qe] * g.det;
    for (int i = tid; i < 2 * N1 * Q1; i += NT) s1[i] = t1d[i];
    for (int i = tid; i < 2 * Q1; i += NT) sh[i] = 0.5 * (i < Q1 ? 1.0 - xq[i] : 1.0 + xq[i - Q1]);
    if constexpr (OVF)
        for (int i = tid; i < 2 * N1 * Q1; i += NT) {
            const int q = i % Q1, bi = (i / Q1) % N1, cx = i / (Q1 * N1);
            sc[i] = 0.5 * (cx ? 1.0 + xq[q] : 1.0 - xq[q]) * t1d[(N1 + bi) * Q1 + q];
        }
    for (int i = tid; i < nee; i += NT) {
        const int d = L[i];
        su[i] = d >= 0 ? u[d] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    for (int it = tid; it < 3 * Q1 * N1; it += NT) {
        const int i = it % N1, qy = (it / N1) % Q1, c = it / (N1 * Q1);
        double2 t1 = make_double2(0.0, 0.0), t2 = make_double2(0.0, 0.0);
#pragma unroll
        for (int j = 0; j < N1; ++j) {
            const double2 a = su[(j * N1 + i) * 3 + c];
            const double vv = s1[j * Q1 + qy], dd = s1[(N1 + j) * Q1 + qy];
            t1.x += vv * a.x; t1.y += vv * a.y;
            t2.x += dd * a.x; t2.y += dd * a.y;
        }
        sT[((c * Q1 + qy) * N1 + i) * 2] = t1;
        sT[((c * Q1 + qy) * N1 + i) * 2 + 1] = t2;
    }
    __syncthreads();
    double2 U[3], G[3][3];
    {
        double vx[N1], dx[N1];
#pragma unroll
        for (int i = 0; i < N1; ++i) { vx[i] = s1[i * Q1 + qx]; dx[i] = s1[(N1 + i) * Q1 + qx]; }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double2 T1 = sT[((c * Q1 + qe) * N1 + ni) * 2], T2 = sT[((c * Q1 + qe) * N1 + ni) * 2 + 1];
            double2 Uc = make_double2(0.0, 0.0), UXI = Uc, UET = Uc;
            if (!ovl) {
                sf_row_sums<NL, 0>(Uc, UXI, UET, T1, T2, vx, dx, std::make_integer_sequence<int, N1>());
            } else {   // the xi = 16 column reads the row's T1, T2 from LDS
                const double2 *r = sT + (c * Q1 + qe) * N1 * 2;
#pragma unroll
                for (int j = 0; j < N1; ++j) {
                    const double2 a1 = r[2 * j], a2 = r[2 * j + 1];
                    Uc.x += vx[j] * a1.x; Uc.y += vx[j] * a1.y;
                    UXI.x += dx[j] * a1.x; UXI.y += dx[j] * a1.y;
                    UET.x += vx[j] * a2.x; UET.y += vx[j] * a2.y;
                }
            }
            G[c][0] = make_double2(rx * UXI.x + sx * UET.x, rx * UXI.y + sx * UET.y);
            G[c][1] = make_double2(kv * Uc.y, -kv * Uc.x);   // d/dy = -i ky
            G[c][2] = make_double2(rz * UXI.x + sz * UET.x, rz * UXI.y + sz * UET.y);
            U[c] = Uc;
        }
    }
    // weighted point values: divergence, symmetrised gradient S_ik = G_ik + G_ki
    // (xx, yy, zz, xy, xz, yz) and -omega^2 u
    const double2 dv = make_double2(w * (G[0][0].x + G[1][1].x + G[2][2].x), w * (G[0][0].y + G[1][1].y + G[2][2].y));
    double2 Sw[6];
    {
        constexpr int si[6] = {0, 1, 2, 0, 0, 1}, sk[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
        for (int j = 0; j < 6; ++j)
            Sw[j] = make_double2(w * (G[si[j]][sk[j]].x + G[sk[j]][si[j]].x), w * (G[si[j]][sk[j]].y + G[sk[j]][si[j]].y));
    }
    const double w2 = -omega * omega * w;
#pragma unroll
    for (int i = 0; i < 3; ++i) U[i] = make_double2(w2 * U[i].x, w2 * U[i].y);
    if (OVF && ovl) {   // the column's point fields, in the order of the row sums below
        const double2 F[17] = {
            make_double2(rx * dv.x, rx * dv.y), make_double2(sx * dv.x, sx * dv.y),
            make_double2(-kv * dv.y, kv * dv.x), make_double2(rz * dv.x, rz * dv.y),
            make_double2(sz * dv.x, sz * dv.y),
            make_double2(rx * Sw[0].x + rz * Sw[4].x, rx * Sw[0].y + rz * Sw[4].y), make_double2(-kv * Sw[3].y, kv * Sw[3].x),
            make_double2(sx * Sw[0].x + sz * Sw[4].x, sx * Sw[0].y + sz * Sw[4].y),
            make_double2(rx * Sw[3].x + rz * Sw[5].x, rx * Sw[3].y + rz * Sw[5].y), make_double2(-kv * Sw[1].y, kv * Sw[1].x),
            make_double2(sx * Sw[3].x + sz * Sw[5].x, sx * Sw[3].y + sz * Sw[5].y),
            make_double2(rx * Sw[4].x + rz * Sw[2].x, rx * Sw[4].y + rz * Sw[2].y), make_double2(-kv * Sw[5].y, kv * Sw[5].x),
            make_double2(sx * Sw[4].x + sz * Sw[2].x, sx * Sw[4].y + sz * Sw[2].y),
            U[0], U[1], U[2]};
#pragma unroll
        for (int f = 0; f < 17; ++f) sF[f * Q1 + qe] = F[f];
    }
    // output slot of this lane in stage 1: (cx, bi), lanes 2 N1.. clamped
    const int slot = min(x16, 2 * N1 - 1), cx = slot / N1, bi = slot % N1;
    double cv[Q1], cdr[OVF ? 1 : Q1];
#pragma unroll
    for (int q = 0; q < Q1; ++q) {
        const double h = sh[cx * Q1 + q];
        cv[q] = h * s1[bi * Q1 + q];
        if constexpr (!OVF) cdr[q] = h * s1[(N1 + bi) * Q1 + q];
    }
    const double *cd = OVF ? sc + (cx * N1 + bi) * Q1 : cdr;
    __syncthreads();   // su and sT are read: sP may overwrite them
    const bool wr = row < Q1 && x16 < 2 * N1;
    const auto seq = std::make_integer_sequence<int, QD>();
    // row sum of point field number f over the xi points of this row
    auto row_sum = [&](double2 &acc, const double2 &F, const double *c, int f) {
        sf_row_dot(acc, F, c, seq);
        if constexpr (OVF) {
            const double2 v = sF[f * Q1 + qe];
            acc.x += c[QD] * v.x;
            acc.y += c[QD] * v.y;
        }
    };
    auto put = [&](int plane, const double2 &v) {
        if (wr) sP[((plane * 2 + cx) * Q1 + qe) * N1 + bi] = v;
    };
    {   // density l: x, y, z rows
        double2 P = make_double2(0.0, 0.0), R = P;
        row_sum(P, make_double2(rx * dv.x, rx * dv.y), cd, 0);
        row_sum(R, make_double2(sx * dv.x, sx * dv.y), cv, 1);
        put(0, P); put(9, R);
        P = make_double2(0.0, 0.0);
        row_sum(P, make_double2(-kv * dv.y, kv * dv.x), cv, 2);
        put(1, P);
        P = make_double2(0.0, 0.0); R = P;
        row_sum(P, make_double2(rz * dv.x, rz * dv.y), cd, 3);
        row_sum(R, make_double2(sz * dv.x, sz * dv.y), cv, 4);
        put(2, P); put(10, R);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {   // density m: rows of component i
        constexpr int ix[3] = {0, 3, 4}, iy[3] = {3, 1, 5}, iz[3] = {4, 5, 2};   // S_ix, S_iy, S_iz in Sw
        const double2 Sx = Sw[ix[i]], Sy = Sw[iy[i]], Sz = Sw[iz[i]];
        double2 P = make_double2(0.0, 0.0), R = P;
        row_sum(P, make_double2(rx * Sx.x + rz * Sz.x, rx * Sx.y + rz * Sz.y), cd, 5 + 3 * i);
        row_sum(P, make_double2(-kv * Sy.y, kv * Sy.x), cv, 6 + 3 * i);
        row_sum(R, make_double2(sx * Sx.x + sz * Sz.x, sx * Sx.y + sz * Sz.y), cv, 7 + 3 * i);
        put(3 + i, P); put(11 + i, R);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {   // rho term
        double2 P = make_double2(0.0, 0.0);
        row_sum(P, U[i], cv, 14 + i);
        put(6 + i, P);
    }
    __syncthreads();
    // contraction over eta: one thread per (corner c, node b, component i)
    for (int it = tid; it < 12 * nen; it += NT) {
        const int i = it % 3, b = (it / 3) % nen, c = it / (3 * nen);
        const int pi = b % N1, pj = b / N1, px = (c == 1 || c == 2) ? 1 : 0, pz = c >= 2 ? 1 : 0;
        const int rl = i == 0 ? 9 : (i == 2 ? 10 : -1);
        double2 l = make_double2(0.0, 0.0), m = l, rr = l;
#pragma unroll 3
        for (int q = 0; q < Q1; ++q) {
            const double h = sh[pz * Q1 + q], fv = h * s1[pj * Q1 + q], fd = h * s1[(N1 + pj) * Q1 + q];
            const double2 *P = sP + (px * Q1 + q) * N1 + pi;   // plane k at P[k * 2 Q1 N1]
            constexpr int PS = 2 * Q1 * N1;
            const double2 a = P[i * PS], b2 = P[(3 + i) * PS], c2 = P[(6 + i) * PS], d2 = P[(11 + i) * PS];
            l.x += fv * a.x; l.y += fv * a.y;
            if (rl >= 0) {
                const double2 f = P[rl * PS];
                l.x += fd * f.x; l.y += fd * f.y;
            }
            m.x += fv * b2.x + fd * d2.x; m.y += fv * b2.y + fd * d2.y;
            rr.x += fv * c2.x; rr.y += fv * c2.y;
        }
        const int a = ie[c] - 1;
        const double r = rho[a], vp = sqrt((lam[a] + 2.0 * mu[a]) / r), vs = sqrt(mu[a] / r);
        const double cl0 = 2.0 * r * vp, cl1 = -4.0 * r * vs, cm1 = 2.0 * r * vs;
        const double cl2 = vp * vp - 2.0 * vs * vs, cm2 = vs * vs;
        const int row3 = b * 3 + i;
        o[((size_t)c * 3 + 0) * nee + row3] = make_double2(cl0 * l.x, cl0 * l.y);
        o[((size_t)c * 3 + 1) * nee + row3] = make_double2(cl1 * l.x + cm1 * m.x, cl1 * l.y + cm1 * m.y);
        o[((size_t)c * 3 + 2) * nee + row3] =
            make_double2(cl2 * l.x + cm2 * m.x + rr.x, cl2 * l.y + cm2 * m.y + rr.y);
    }
}

// Pseudo-Hessian diagonal, anchor part: the column of anchor a is the sum of
// its elements' vectors on shared DOFs.  pmap[(a*ncon + s)*nee + j] is the
// patch position of row j of the s-th element of a (-1 fixed DOF); the
// element and its corner come from mconn.  diag[3a + p] += |column|^2.
__global__ __launch_bounds__(256) void k_fdiag(int nnpg, int ncon, int nee, const int *__restrict__ mconn,
                                                const int *__restrict__ ieng, const int *__restrict__ pmap,
                                                const char *__restrict__ fixa, const double2 *__restrict__ cols,
                                                double *__restrict__ diag) {
    extern __shared__ double sm[];
    __shared__ double red[3][16];
    const int a = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    if (a >= nnpg || fixa[a]) return;   // anchors of PML elements: fixed parameters
    const int np = ncon * nee;
    double2 *acc = reinterpret_cast<double2 *>(sm);   // np * 3
    for (int i = tid; i < np * 3; i += nt) acc[i] = make_double2(0.0, 0.0);
    __syncthreads();
    for (int s = 0; s < ncon; ++s) {
        const int e1 = mconn[(size_t)s * nnpg + a];
        if (e1 <= 0) break;
        const int e = e1 - 1;
        int corner = 0;
        for (int k = 0; k < 4; ++k)
            if (ieng[4 * e + k] - 1 == a) corner = k;
        const int *pm = pmap + ((size_t)a * ncon + s) * nee;
        const double2 *src = cols + ((size_t)e * 4 + corner) * 3 * nee;
        // one thread per (parameter p, row j): the rows of one element map
        // to distinct positions, so no two threads add to the same value
        for (int i = tid; i < 3 * nee; i += nt) {
            const int p = i / nee, j = i - p * nee;
            const int pos = pm[j];
            if (pos < 0) continue;
            const double2 v = src[i];
            acc[pos * 3 + p].x += v.x;
            acc[pos * 3 + p].y += v.y;
        }
        __syncthreads();
    }
    double s3[3] = {0.0, 0.0, 0.0};
    for (int i = tid; i < np; i += nt)
        for (int p = 0; p < 3; ++p) s3[p] += acc[i * 3 + p].x * acc[i * 3 + p].x + acc[i * 3 + p].y * acc[i * 3 + p].y;
#pragma unroll
    for (int p = 0; p < 3; ++p)
        for (int o = 32; o > 0; o >>= 1) s3[p] += __shfl_xor(s3[p], o);
    if ((tid & 63) == 0)
        for (int p = 0; p < 3; ++p) red[p][tid >> 6] = s3[p];
    __syncthreads();
    if (tid < 3) {
        double t = 0.0;
        for (int w = 0; w < (nt + 63) / 64; ++w) t += red[tid][w];
        diag[3 * a + tid] += t;
    }
}

// ADJFLAM_DIST: gprod(mygrad(c)) -= Re(conj(F) lambda) over column c
__global__ void k_adjflam(int nnpgl, const int *__restrict__ mygrad, const int *__restrict__ jcsc,
                          const int *__restrict__ icsc, const float2 *__restrict__ f,
                          const float2 *__restrict__ lam, float *__restrict__ g) {
    const int c = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    const int l = threadIdx.x & 63;
    if (c >= nnpgl) return;
    double acc = 0.0;
    for (int i = jcsc[c] - 1 + l; i < jcsc[c + 1] - 1; i += 64) {
        float2 a = f[i], b = lam[icsc[i] - 1];
        acc += (double)(a.x * b.x + a.y * b.y);
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (l == 0) g[mygrad[c] - 1] -= (float)acc;
}

// ---------------------------------------------------------------------------
// Element derivative products, collocated on one anchor (per column) or
// summed over anchors (per element):
//  * DABMKISOB25 (dstiff25.f:1-203), the second-order products of QPMFREE
//    (qpmfree.f90:622-1215): IMAG, ky coupling in the imaginary part;
//  * DKM_VPVS (dstiff25.f:498-662) for CFMAT_DIST/VFORCE25 (hessloc.f90,
//    vforce25.f): !IMAG, ky coupling written as real numbers.
// Both are the elastic element operator without the -w^2 rho mass term,
// with density-normalised moduli at the quadrature points
//   lambda + 2 mu = P = kp D,  mu = S = ks Ds,  lambda = Q = P - 2 S,
//   D(q) = w(q) sum_k c_k N_k(q),  Ds(q) = w(q) sum_k s_k N_k(q)
// (DABMKISOB25: kp, ks = 1, 0 for Vp, 0, 1 for Vs, with s_k = c_k, and
// 1, 1 for the joint P-S inversion, whose P and S variables carry their own
// corner coefficients c_k and s_k; DKM_VPVS: the VPVS slaving of
// vforce25.f:164-179, s_k = c_k), eight coefficients per item.
// The operator M is applied to nvec <= 2 vectors through fluxes at the
// quadrature points:
//   mode 0 (item = element; one launch per colour, so rows never collide):
//          y_v += M x_v (M is Hermitian: M x is also M^H x);
//   mode 1 (item = local column c, its element the last connected one):
//          p_v(MYGRAD(c)) = -sum over the element's rows Re(conj(M x_v) z_v);
//   mode 2 (item = local column c): column c of F in CSC order, COMPLEX*8,
//          zero outside the element's rows.
// The rows of modes 1 and 2 come with their positions in column c's CSC
// pattern (rpos[c][row], built and checked on the host once per inversion
// structure: a per-row search of the pattern in the kernel was a chain of
// dependent loads).
// Fluxes per (vector, q), 8 complex: comp 1 N, Nx, Nz; comp 2 N, Nx, Nz;
// comp 3 N, Nz (comp 3's Nx flux equals comp 1's Nz flux).
// Row outputs of one item (rows are threads 0..nee-1, row = -1 for
// the others): mode 0 adds into y, mode 1 contracts with z into p(MYGRAD),
// mode 2 writes column it of F at the rows' CSC positions (err 3 if the
// table has none).
__device__ inline void dab_epilogue(int mode, int nvec, int it, int row, const int *__restrict__ L,
                                    const double2 *val, const int *__restrict__ rpos,
                                    const int *__restrict__ mygrad, double2 *__restrict__ y0,
                                    double2 *__restrict__ y1, const float2 *__restrict__ z0,
                                    const float2 *__restrict__ z1, float *__restrict__ p0, float *__restrict__ p1,
                                    float2 *__restrict__ fout, int *__restrict__ err, double (*red)[8]) {
    double acc[2] = {0.0, 0.0};
    const int dof = row >= 0 ? L[row] : -1;
    if (dof >= 0) {
        if (mode == 0) {
            y0[dof].x += val[0].x; y0[dof].y += val[0].y;
            if (nvec > 1) { y1[dof].x += val[1].x; y1[dof].y += val[1].y; }
        } else {
            // the row's position in column it's CSC pattern, from the host's
            // table (rpos: this item's rows), which has checked that every
            // row is in the pattern
            const int pos = rpos[row];
            if (pos < 0) {
                *err = 3;
            } else if (mode == 2) {
                fout[pos] = make_float2((float)val[0].x, (float)val[0].y);
            } else {
                const float2 za = z0[dof];
                acc[0] -= val[0].x * za.x + val[0].y * za.y;
                if (nvec > 1) {
                    const float2 zb = z1[dof];
                    acc[1] -= val[1].x * zb.x + val[1].y * zb.y;
                }
            }
        }
    }
    if (mode != 1) return;
#pragma unroll
    for (int v = 0; v < 2; ++v)
        for (int o = 32; o > 0; o >>= 1) acc[v] += __shfl_xor(acc[v], o);
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { red[0][wv] = acc[0]; red[1][wv] = acc[1]; }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = blockDim.x >> 6;
        double s0 = 0.0, s1 = 0.0;
        for (int w = 0; w < nw; ++w) { s0 += red[0][w]; s1 += red[1][w]; }
        p0[mygrad[it] - 1] = (float)s0;
        if (nvec > 1) p1[mygrad[it] - 1] = (float)s1;
    }
}

// The element derivative products above by sum factorisation (every
// order, p2..p6), one workgroup per item.  Per vector, the element field and its
// reference derivatives at the quadrature points come from two 1-D passes
// (as in k_grad_sf: T1, T2 along eta, then rows of 16 lanes along xi with
// the DPP broadcast), and the projection on the test functions is split the
// same way: with A = rx F_x + rz F_z, B = sx F_x + sz F_z, C = F_N per
// component at every point, row (j, i) of the element vector is
//   sum_eta v_j(eta) (PA + PC)[eta][i] + v'_j(eta) PB[eta][i],
//   PA = sum_xi v'_i(xi) A,  PB = sum_xi v_i(xi) B,  PC = sum_xi v_i(xi) C.
// About 4x fewer multiply-adds than the direct sums over the element
// nodes and the quadrature points.  The 1-D tables v, v' sit in LDS, which
// keeps the kernel at 120 registers: four waves per SIMD, so two five-wave
// workgroups fit per CU at p6 (a workgroup holds ceil(waves / 4) slots on
// every SIMD; at three waves per SIMD one fits) and four four-wave ones at
// p4.
template <int NL, bool IMAG>
__global__ __launch_bounds__((2 * NL + 3 > 16 ? 320 : 256)) __attribute__((amdgpu_waves_per_eu(4))) void k_dab_sf(int mode, int nitems, int nvec, double kp, double ks,
                                                double kv, const int *__restrict__ item_elem,
                                                const double *__restrict__ item_coef,
                                                const double *__restrict__ t1d, const double *__restrict__ wq,
                                                const double *__restrict__ xq, const int *__restrict__ ieng,
                                                const double *__restrict__ xl, const double *__restrict__ zl,
                                                const int *__restrict__ lm, const float2 *__restrict__ x0,
                                                const float2 *__restrict__ x1, double2 *__restrict__ y0,
                                                double2 *__restrict__ y1, const float2 *__restrict__ z0,
                                                const float2 *__restrict__ z1, const int *__restrict__ mygrad,
                                                const int *__restrict__ jcsc, const int *__restrict__ rpos,
                                                float *__restrict__ p0, float *__restrict__ p1,
                                                float2 *__restrict__ fout, int *__restrict__ err) {
    constexpr int N1 = NL, Q1 = 2 * NL + 3, nen = NL * NL, nee = 3 * nen, QQ = Q1 * Q1;
    // Q1 = 17 (p6): rows 0..16 of 16 lanes hold xi points 0..15 and the 17
    // lanes after them xi point 16 of eta rows 0..16 (five waves)
    constexpr bool OVF = Q1 > 16;
    constexpr int NT = OVF ? 320 : 256;
    // first-pass and projection items: (c, eta point, i) when 3 Q1 N1 fit
    // the threads, else (c, eta pair qh and qh + QH, i)
    constexpr int QH = 3 * Q1 * N1 <= NT ? Q1 : (Q1 + 1) / 2;
    static_assert(3 * QH * N1 <= NT, "one round of first-pass and projection items");
    static_assert(Q1 <= 17, "one row of 16 lanes per eta point, plus one column");
    extern __shared__ double sm[];
    __shared__ double red[2][8];
    double2 *su = reinterpret_cast<double2 *>(sm);          // [2][nee]
    double2 *sT = su + 2 * nee;                             // [3][Q1][N1][2]: T1, T2
    // PA, PB, PC [3][Q1][N1][3] take T1/T2's place: T1/T2 are last read in
    // the point stage, PA.. first written in the projection after it
    double2 *sP = sT;
    double2 *sH = sT + 3 * Q1 * N1 * 3;                     // [3 comp][3: A, B, C][QQ]
    double *sb = reinterpret_cast<double *>(sH + 9 * QQ);    // [2][N1][Q1]: v_k(q), v'_k(q)
    const int it = blockIdx.x, tid = threadIdx.x;
    if (it >= nitems) return;
    const int e = item_elem[it];
    // corner coefficients of the P field (0..3) and of the S field (4..7)
    const double *ic = item_coef + 8 * (size_t)it;
    const double c0 = ic[0], c1 = ic[1], c2 = ic[2], c3 = ic[3], s0 = ic[4], s1 = ic[5], s2 = ic[6], s3 = ic[7];
    const int *ie = ieng + 4 * e;
    const int *L = lm + (size_t)e * nee;
    for (int i = tid; i < 2 * N1 * Q1; i += NT) sb[i] = t1d[i];
    const double *v1 = sb, *d1 = sb + N1 * Q1;   // v1[k * Q1 + q]
    if (mode == 2)
        for (int r = jcsc[it] - 1 + tid; r < jcsc[it + 1] - 1; r += NT) fout[r] = make_float2(0.f, 0.f);
    for (int i = tid; i < nee; i += NT) {
        const int d = L[i];
        float2 a = d >= 0 ? x0[d] : make_float2(0.f, 0.f);
        su[i] = make_double2(a.x, a.y);
        if (nvec > 1) {
            float2 b = d >= 0 ? x1[d] : make_float2(0.f, 0.f);
            su[nee + i] = make_double2(b.x, b.y);
        }
    }
    // stage-2 lanes: row = eta point, lane = xi point (clamped when idle)
    const int row16 = tid >> 4, x16 = tid & 15;
    const bool ovl = OVF && tid >= 16 * Q1 && tid < 17 * Q1;   // a lane of the xi = 16 column
    const bool act = (row16 < Q1 && x16 < Q1) || ovl;
    const int qe = ovl ? tid - 16 * Q1 : min(row16, Q1 - 1), qx = ovl ? 16 : min(x16, Q1 - 1);
    // the inverse Jacobian and the collocated density D at this lane's point,
    // in registers (an LDS table of the points would keep fewer workgroups
    // per CU)
    double rx, sx, rz, sz, D, Ds;
    {
        Geo g;
        if (geo_at(xq[qx], xq[qe], ie, xl, zl, g)) *err = 2;
        rx = g.dzdeta / g.det; sx = -g.dzdxi / g.det; rz = -g.dxdeta / g.det; sz = g.dxdxi / g.det;
        D = wq[qx] * wq[qe] * g.det * (c0 * g.sf[0] + c1 * g.sf[1] + c2 * g.sf[2] + c3 * g.sf[3]);
        Ds = wq[qx] * wq[qe] * g.det * (s0 * g.sf[0] + s1 * g.sf[1] + s2 * g.sf[2] + s3 * g.sf[3]);
    }
    const int ni = min(x16, N1 - 1);
    const double k2 = kv * kv;
    // output row of this thread: node (j, i) = (a / N1, a % N1), component cc
    const int orow = tid < nee ? tid : -1;
    const int oa = tid / 3, occ = tid % 3, oj = oa / N1, oi = oa % N1;
    double2 val[2] = {};
    for (int v = 0; v < nvec; ++v) {
        __syncthreads();   // su written; the previous vector is done with sT, sH, sP
        // one item per (c, eta point pair qh, qh + QH, i): a single round of
        // the workgroup's threads also at p6 (3 Q1 N1 = 357 > 320)
        for (int it = tid; it < 3 * QH * N1; it += NT) {
            const int i = it % N1, qh = (it / N1) % QH, c = it / (N1 * QH);
            double2 a[N1];
#pragma unroll
            for (int j = 0; j < N1; ++j) a[j] = su[v * nee + (j * N1 + i) * 3 + c];
#pragma unroll
            for (int h = 0; h < Q1 / QH + (Q1 % QH ? 1 : 0); ++h) {
                const int qy = qh + h * QH;
                if (qy >= Q1) break;
                double2 t1 = make_double2(0.0, 0.0), t2 = t1;
#pragma unroll
                for (int j = 0; j < N1; ++j) {
                    const double vv = v1[j * Q1 + qy], dd = d1[j * Q1 + qy];
                    t1.x += vv * a[j].x; t1.y += vv * a[j].y;
                    t2.x += dd * a[j].x; t2.y += dd * a[j].y;
                }
                const int kk = (c * Q1 + qy) * N1 + i;
                sT[kk * 2] = t1;
                sT[kk * 2 + 1] = t2;
            }
        }
        __syncthreads();
        {
            const double P = kp * D, S = ks * Ds, Q = P - 2.0 * S;
            double2 U[3], UX[3], UZ[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double2 *o = sT + ((c * Q1 + qe) * N1 + ni) * 2;
                const double2 T1 = o[0], T2 = o[1];
                double2 u0 = make_double2(0.0, 0.0), uxi = u0, uet = u0;
                if (!ovl) {
                    sf_row_sums_lds<NL, Q1>(u0, uxi, uet, T1, T2, v1 + qx, std::make_integer_sequence<int, N1>());
                } else {   // the xi = 16 column reads the row's T1, T2 from LDS
                    const double2 *r = o - 2 * ni;
#pragma unroll
                    for (int j = 0; j < N1; ++j) {
                        const double2 a1 = r[2 * j], a2 = r[2 * j + 1];
                        const double vj = v1[j * Q1 + qx], dj = d1[j * Q1 + qx];
                        u0.x += vj * a1.x; u0.y += vj * a1.y;
                        uxi.x += dj * a1.x; uxi.y += dj * a1.y;
                        uet.x += vj * a2.x; uet.y += vj * a2.y;
                    }
                }
                U[c] = u0;
                UX[c] = make_double2(rx * uxi.x + sx * uet.x, rx * uxi.y + sx * uet.y);
                UZ[c] = make_double2(rz * uxi.x + sz * uet.x, rz * uxi.y + sz * uet.y);
            }
            auto mk = [&](double2 X) {
                return IMAG ? make_double2(-kv * X.y, kv * X.x) : make_double2(kv * X.x, kv * X.y);
            };
            auto lin = [](double a, double2 X, double b, double2 Y) {
                return make_double2(a * X.x + b * Y.x, a * X.y + b * Y.y);
            };
            // the fluxes above
            const double2 m2 = mk(U[1]);
            double2 F[8];
            F[0] = lin(S, mk(UX[1]), k2 * S, U[0]);
            F[1] = make_double2(P * UX[0].x + Q * UZ[2].x - Q * m2.x, P * UX[0].y + Q * UZ[2].y - Q * m2.y);
            F[2] = make_double2(S * (UZ[0].x + UX[2].x), S * (UZ[0].y + UX[2].y));
            F[3] = lin(Q, mk(make_double2(UX[0].x + UZ[2].x, UX[0].y + UZ[2].y)), k2 * P, U[1]);
            F[4] = lin(S, UX[1], -S, mk(U[0]));
            F[5] = lin(S, UZ[1], -S, mk(U[2]));
            F[6] = lin(S, mk(UZ[1]), k2 * S, U[2]);
            F[7] = make_double2(Q * UX[0].x + P * UZ[2].x - Q * m2.x, Q * UX[0].y + P * UZ[2].y - Q * m2.y);
            if (act) {
                const int q = qx + Q1 * qe;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const double2 fN = F[3 * c], fX = F[c == 2 ? 2 : 3 * c + 1], fZ = F[c == 2 ? 7 : 3 * c + 2];
                    sH[(c * 3 + 0) * QQ + q] = lin(rx, fX, rz, fZ);
                    sH[(c * 3 + 1) * QQ + q] = lin(sx, fX, sz, fZ);
                    sH[(c * 3 + 2) * QQ + q] = fN;
                }
            }
        }
        __syncthreads();
        for (int it = tid; it < 3 * QH * N1; it += NT) {
            const int i = it % N1, qh = (it / N1) % QH, c = it / (N1 * QH);
            constexpr int NH = Q1 / QH + (Q1 % QH ? 1 : 0);
            double2 pa[NH], pb[NH], pc[NH];
#pragma unroll
            for (int h = 0; h < NH; ++h) pa[h] = pb[h] = pc[h] = make_double2(0.0, 0.0);
            auto xstep = [&](int x) {
                const double vv = v1[i * Q1 + x], dd = d1[i * Q1 + x];
#pragma unroll
                for (int h = 0; h < NH; ++h) {
                    const int qy = min(qh + h * QH, Q1 - 1);
                    const double2 A = sH[(c * 3 + 0) * QQ + qy * Q1 + x], B = sH[(c * 3 + 1) * QQ + qy * Q1 + x],
                                  Cc = sH[(c * 3 + 2) * QQ + qy * Q1 + x];
                    pa[h].x += dd * A.x; pa[h].y += dd * A.y;
                    pb[h].x += vv * B.x; pb[h].y += vv * B.y;
                    pc[h].x += vv * Cc.x; pc[h].y += vv * Cc.y;
                }
            };
            if constexpr (NH > 1) {   // two eta points: a rolled loop keeps the registers at 127
#pragma unroll 1
                for (int x = 0; x < Q1; ++x) xstep(x);
            } else {
                for (int x = 0; x < Q1; ++x) xstep(x);
            }
#pragma unroll
            for (int h = 0; h < NH; ++h) {
                const int qy = qh + h * QH;
                if (qy >= Q1) break;
                const int kk = (c * Q1 + qy) * N1 + i;
                sP[kk * 3] = pa[h];
                sP[kk * 3 + 1] = pb[h];
                sP[kk * 3 + 2] = pc[h];
            }
        }
        __syncthreads();
        if (orow >= 0) {
            double2 r = make_double2(0.0, 0.0);
            for (int qy = 0; qy < Q1; ++qy) {
                const double2 *o = sP + ((occ * Q1 + qy) * N1 + oi) * 3;
                const double vv = v1[oj * Q1 + qy], dd = d1[oj * Q1 + qy];
                r.x += vv * (o[0].x + o[2].x) + dd * o[1].x;
                r.y += vv * (o[0].y + o[2].y) + dd * o[1].y;
            }
            if (v == 0) val[0] = r; else val[1] = r;
        }
    }
    dab_epilogue(mode, nvec, it, orow, L, val, mode ? rpos + (size_t)it * nee : nullptr, mygrad, y0, y1, z0, z1, p0,
                 p1, fout, err, red);
}

// ---------------------------------------------------------------------------
// Jacobian of one (frequency, source): J(irhs, c) = sum over column c of F
// of G(ICSC(i), irhs) F(i)  (PCHADJ's product, pchadj.f90:239-254), with
// G transposed to [dof][rhs] so a wave reads one contiguous row of G per
// nonzero of F.  One wave per column, lanes over the right-hand sides,
// FP64 accumulation; hdiag(MYGRAD(c)) += sum_irhs |J|^2 (the 1x1 CGEMM of
// pchadj.f90:264-281).  jac[c*nrhs + irhs].
__global__ __launch_bounds__(256) void k_pchadj(int nnpgl, int nrhs, const int *__restrict__ mygrad,
                                                 const int *__restrict__ jcsc, const int *__restrict__ icsc,
                                                 const float2 *__restrict__ gt, const float2 *__restrict__ f,
                                                 float2 *__restrict__ jac, float *__restrict__ hdiag) {
    // a lane owns right-hand sides r0 + lane + 64 j (j < 3) of a 192-wide
    // pass.  The column's nonzeros are read 64 at a time, one per lane, with
    // their rows (coalesced); a ballot keeps those whose F entry is not
    // exactly zero (a column of CFMAT_DIST holds only its last element's
    // rows, about a third of the pattern: the sums are unchanged for finite
    // G), and the wave walks them two at a time in pattern order, their
    // rows and values broadcast from the lane that read them: six rows of G
    // in flight per lane, no dependent scalar loads in the loop, 62
    // registers (eight waves per SIMD; four at a time took 98 registers and
    // ran 5.20 against 3.50 ms, eight 6.54 ms: 400x200 p6, 180 right-hand
    // sides).
    constexpr int R = 3, U = 2;
    const int c = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    const int lane = threadIdx.x & 63;
    if (c >= nnpgl) return;
    const int i1 = jcsc[c] - 1, i2 = jcsc[c + 1] - 1;
    double hs = 0.0;
    for (int r0 = 0; r0 < nrhs; r0 += 64 * R) {
        double ar[R], ai[R];
        bool ok[R];
#pragma unroll
        for (int j = 0; j < R; ++j) { ar[j] = ai[j] = 0.0; ok[j] = r0 + lane + 64 * j < nrhs; }
        for (int base = i1; base < i2; base += 64) {
            const int i = base + lane;
            float2 fl = make_float2(0.f, 0.f);
            int rl = -1;
            if (i < i2) {
                fl = f[i];
                if (fl.x != 0.f || fl.y != 0.f) rl = icsc[i] - 1;
            }
            unsigned long long m = __ballot(rl >= 0);
            while (m) {
                int ru[U];
                float fx[U], fy[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    ru[u] = -1;
                    fx[u] = fy[u] = 0.f;
                    if (m) {
                        const int k = __ffsll((long long)m) - 1;
                        m &= m - 1;
                        ru[u] = __builtin_amdgcn_readlane(rl, k);
                        fx[u] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(fl.x), k));
                        fy[u] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(fl.y), k));
                    }
                }
                float2 g[U][R];
#pragma unroll
                for (int u = 0; u < U; ++u)
#pragma unroll
                    for (int j = 0; j < R; ++j)
                        g[u][j] = (ok[j] && ru[u] >= 0) ? gt[(size_t)ru[u] * nrhs + r0 + lane + 64 * j]
                                                        : make_float2(0.f, 0.f);
#pragma unroll
                for (int u = 0; u < U; ++u)
#pragma unroll
                    for (int j = 0; j < R; ++j) {
                        if (ru[u] < 0) continue;
                        ar[j] += (double)g[u][j].x * fx[u] - (double)g[u][j].y * fy[u];
                        ai[j] += (double)g[u][j].x * fy[u] + (double)g[u][j].y * fx[u];
                    }
            }
        }
#pragma unroll
        for (int j = 0; j < R; ++j) {
            if (!ok[j]) continue;
            const float2 jv = make_float2((float)ar[j], (float)ai[j]);
            jac[(size_t)c * nrhs + r0 + lane + 64 * j] = jv;
            hs += (double)jv.x * jv.x + (double)jv.y * jv.y;
        }
    }
    for (int o = 32; o > 0; o >>= 1) hs += __shfl_xor(hs, o);
    if (lane == 0) hdiag[mygrad[c] - 1] += (float)hs;
}

// GNMFREE's products (gnmfree.f90:185-212) over a batch of Jacobians with
// uneven observation counts, J_j stored column-major (nobs_j x nnpgl) at
// jac + joff[j]:  JP_j(obs) = sum_c J_j(obs, c) p(MYGRAD(c))  (FP64).
// Grid (column chunks, j); block (ch, j) sums chunk ch of the columns for
// every observation of J_j into part[ch * pt + poff[j] + obs].  k_gn_jsum
// then adds the chunks in chunk order, so JP does not depend on the order
// in which blocks run (no atomics: bit-identical results run to run).
__global__ void k_gn_jp(int nnpgl, int cch, int64_t pt, const int *__restrict__ nobs,
                        const int64_t *__restrict__ joff, const int64_t *__restrict__ poff,
                        const int *__restrict__ mygrad, const float2 *__restrict__ jac,
                        const double *__restrict__ p, double2 *__restrict__ part) {
    const int j = blockIdx.y;
    const int no = nobs[j];
    const int c0 = blockIdx.x * cch, c1 = min(nnpgl, c0 + cch);
    const float2 *J = jac + joff[j];
    double2 *out = part + (size_t)blockIdx.x * pt + poff[j];
    for (int o = threadIdx.x; o < no; o += blockDim.x) {
        double ar = 0.0, ai = 0.0;
        // eight columns per step: eight loads in flight per lane
        int c = c0;
        for (; c + 8 <= c1; c += 8) {
            float2 v[8];
            double pv[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                v[k] = J[(size_t)(c + k) * no + o];
                pv[k] = p[mygrad[c + k] - 1];
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                ar += (double)v[k].x * pv[k];
                ai += (double)v[k].y * pv[k];
            }
        }
        for (; c < c1; ++c) {
            const double pv = p[mygrad[c] - 1];
            const float2 v = J[(size_t)c * no + o];
            ar += (double)v.x * pv;
            ai += (double)v.y * pv;
        }
        out[o] = make_double2(ar, ai);
    }
}

// jp[o] = sum over chunks ch = 0..nch-1 of part[ch * pt + o], in chunk order.
__global__ void k_gn_jsum(int nch, int64_t pt, const double2 *__restrict__ part, double2 *__restrict__ jp) {
    for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < pt; o += (int64_t)gridDim.x * blockDim.x) {
        double ar = 0.0, ai = 0.0;
        for (int ch = 0; ch < nch; ++ch) {
            const double2 v = part[(size_t)ch * pt + o];
            ar += v.x;
            ai += v.y;
        }
        jp[o] = make_double2(ar, ai);
    }
}

// z(MYGRAD(c)) += sum_j alpha_j Re( sum_obs conj(J_j(obs, c)) JP_j(obs) ),
// one wave per column.
__global__ void k_gn_z(int nnpgl, int njac, const int *__restrict__ nobs, const int64_t *__restrict__ joff,
                       const int64_t *__restrict__ poff, const double *__restrict__ alpha,
                       const int *__restrict__ mygrad, const float2 *__restrict__ jac,
                       const double2 *__restrict__ jp, double *__restrict__ z) {
    const int c = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    const int lane = threadIdx.x & 63;
    if (c >= nnpgl) return;
    double s = 0.0;
    // four Jacobians per step: their column loads are in flight together
    int j = 0;
    for (; j + 4 <= njac; j += 4) {
        double sj[4] = {0.0, 0.0, 0.0, 0.0};
        int no[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) no[k] = nobs[j + k];
        const int nmax = max(max(no[0], no[1]), max(no[2], no[3]));
        for (int o = lane; o < nmax; o += 64) {
            float2 v[4];
            double2 a[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const bool in = o < no[k];
                v[k] = in ? jac[joff[j + k] + (size_t)c * no[k] + o] : make_float2(0.f, 0.f);
                a[k] = in ? jp[poff[j + k] + o] : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) sj[k] += (double)v[k].x * a[k].x + (double)v[k].y * a[k].y;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) s += alpha[j + k] * sj[k];
    }
    for (; j < njac; ++j) {
        const int no = nobs[j];
        const float2 *J = jac + joff[j] + (size_t)c * no;
        const double2 *P = jp + poff[j];
        double sj = 0.0;
        for (int o = lane; o < no; o += 64) {
            const float2 v = J[o];
            const double2 a = P[o];
            sj += (double)v.x * a.x + (double)v.y * a.y;
        }
        s += alpha[j] * sj;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) z[mygrad[c] - 1] += s;
}

// Rows of the Gauss-Newton Hessian's factor for Jacobian j (nobs x nnpgl
// COMPLEX*8, column-major): the real parts in rows roff .. roff+no-1 and the
// imaginary parts in the next no rows of the FP64 column-major matrix A
// (leading dimension lda), scaled by sa = sqrt(alpha_j).
__global__ void k_jac_rows(int nnpgl, int no, double sa, const float2 *__restrict__ jac, double *__restrict__ a,
                           int64_t lda, int64_t roff) {
    const int64_t n = (int64_t)no * nnpgl;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = i / no, o = i - c * no;
        const float2 v = jac[i];
        a[c * lda + roff + o] = sa * (double)v.x;
        a[c * lda + roff + no + o] = sa * (double)v.y;
    }
}

// H(j, i) = H(i, j) for i < j (column-major n x n): the lower triangle from
// the upper one.
__global__ void k_sym_lower(int n, double *__restrict__ h) {
    const int64_t tot = (int64_t)n * n;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < tot; k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = k / n, i = k - j * n;   // element (i, j)
        if (i > j) h[k] = h[j + i * (int64_t)n];
    }
}

// ---------------------------------------------------------------------------
// Adjoint-source pipeline (the vector work between a forward solve and the
// adjoint solve): receiver extraction (EXRESP25D, locrec2d.f:867-925) and
// CADJLAM's right-hand side (adjgrad.f90:229-300: rotation to (N,E,Z),
// receiver phase shift, BPRESID4 residuals, rotation back, RESID_DRHS).
// COMPLEX*8 arithmetic in the reference's forms: products and quotients by
// the textbook formulas.  |z|, sin, cos and atan2 of single-precision
// arguments are evaluated in double and rounded once, which reproduces the
// host library's single-precision results (rounded to within half an ulp)
// far more often than the device's own single-precision routines: the L1
// residual dif/|dif| magnifies any difference by 1/|dif|.
__device__ inline float2 cmulf(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline float2 cdivf(float2 a, float2 b) {
    const float d = b.x * b.x + b.y * b.y;
    return make_float2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
__device__ inline float cabsf_(float2 a) {
    return (float)sqrt((double)a.x * a.x + (double)a.y * a.y);
}
__device__ inline float2 cexpi(float t) {   // exp(i t)
    double s, c;
    sincos((double)t, &s, &c);
    return make_float2((float)c, (float)s);
}
__device__ inline float atan2r(float y, float x) { return (float)atan2((double)y, (double)x); }
__device__ inline void crotate(float theta, float2 q1, float2 q2, float2 &r1, float2 &r2) {
    const float pi180 = 0.017453292519943295f;
    const float2 cs = cexpi(theta * pi180);
    const float2 c = make_float2(cs.x, 0.f), s = make_float2(cs.y, 0.f);
    const float2 a = cmulf(q1, c), b = cmulf(q2, s), e = cmulf(q1, s), d = cmulf(q2, c);
    r1 = make_float2(a.x - b.x, a.y - b.y);
    r2 = make_float2(e.x + d.x, e.y + d.y);
}
__device__ inline float2 cphm2cm(float mag, float ph) { return cmulf(make_float2(mag, 0.f), cexpi(ph)); }

// objfn.f:972-1009 CRESID
__device__ inline float2 cresid(int irestp, float2 obs, float2 est) {
    if (obs.x == 0.f && obs.y == 0.f) return make_float2(0.f, 0.f);
    if (irestp == 1 || irestp == 2) {
        float ad, pd, ae, pe;
        if (irestp == 2) {
            ad = sqrtf(obs.x * obs.x + obs.y * obs.y); pd = 0.f;
            ae = sqrtf(est.x * est.x + est.y * est.y); pe = 0.f;
        } else {
            ad = 1.f; pd = atan2r(obs.y, obs.x);
            ae = 1.f; pe = atan2r(est.y, est.x);
        }
        const float2 a = cmulf(make_float2(ad, 0.f), cexpi(pd)), b = cmulf(make_float2(ae, 0.f), cexpi(pe));
        return make_float2(a.x - b.x, a.y - b.y);
    }
    return make_float2(obs.x - est.x, obs.y - est.y);
}

// objfn.f:757-839 BPRESID4 for one component (EPS = 0 as in the reference)
__device__ inline float2 bpresid(int norm, int irestp, float epss, double freq, float dtmax, float w, float2 obs,
                                 float2 est) {
    const float twopi = 6.283185307179586f, pi = 3.1415926535897931f, eps = 0.f, eps2 = epss * (eps * eps);
    if (cabsf_(obs) == 0.f) return make_float2(0.f, 0.f);
    const float2 dif = cresid(irestp, obs, est), wt = make_float2(w, 0.f), wc = make_float2(w, -0.f);
    if (norm == 1) return cdivf(cmulf(wc, dif), make_float2(cabsf_(dif), 0.f));
    if (norm == 3) {
        if (cabsf_(cmulf(wt, dif)) <= eps) {
            const float aw = cabsf_(wt);
            return cmulf(make_float2(aw * aw / eps, 0.f), dif);
        }
        return cdivf(cmulf(wc, dif), make_float2(cabsf_(dif), 0.f));
    }
    const float aw = cabsf_(wt);
    if (norm == 4) {
        const float a = cabsf_(cmulf(wt, dif));
        const float den = eps2 * sqrtf(1.f + a * a / eps2);
        return cmulf(make_float2(aw * aw / den, 0.f), dif);
    }
    float2 r = cmulf(make_float2(aw * aw, 0.f), dif);
    if (dtmax > 0.f) {
        float dphi = atan2r(obs.y, obs.x) - atan2r(est.y, est.x);
        if (dphi < -pi) dphi = dphi + twopi;
        if (dphi > +pi) dphi = dphi - twopi;
        if (fabs((double)dphi / ((double)twopi * freq)) > (double)dtmax) r = make_float2(0.f, 0.f);
    }
    return r;
}

// adjgrad.f90:527-566 HDIFFER (sqrt(-i w) = sqrt(w/2)(1 - i), ...)
__device__ inline float2 hdiffer(int ib, double freq) {
    const float om = (float)(2.0 * 3.1415926535897932384626434 * freq);
    const float s = sqrtf(0.5f * om);
    switch (ib) {
        case 1: return make_float2(s, -s);
        case 2: return make_float2(-s, s);
        case 3: return make_float2(s, s);
        case 4: return make_float2(-s, -s);
        case 5: return make_float2(0.f, om);
        case 6: return make_float2(0.f, 1.f);
        default: return make_float2(1.f, 0.f);
    }
}

// EST(i, irec, isrc) = RECV(i, irec) SOL(isrc, MRDOF(i, irec)) exp(+i w PY(isrc) YREC(irec)),
// one thread per (source, receiver); MRDOF in 0..NDOF (checked by the host)
__global__ void k_exresp(int ndof, int nsrc, int nrec, double omega, const double *__restrict__ py,
                         const double *__restrict__ yrec, const float2 *__restrict__ sol,
                         const int *__restrict__ mrdof, const float2 *__restrict__ recv, float2 *__restrict__ est) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nsrc * nrec) return;
    const int is = t / nrec, ir = t % nrec;
    const float2 cf = cexpi((float)(omega * py[is] * yrec[ir]));
    for (int i = 0; i < 3; ++i) {
        const int d = mrdof[3 * ir + i];
        float2 v = make_float2(0.f, 0.f);
        if (d != 0) v = cmulf(cmulf(recv[3 * ir + i], sol[(size_t)is * ndof + d - 1]), cf);
        est[3 * ((size_t)nrec * is + ir) + i] = v;
    }
}

struct AdjParams {
    float az, epss, dtmax;
    int norm, irestp, ibphase, lunwrap;
};

// CADJLAM's right-hand side for npair (ifreq, isrc) pairs, one workgroup
// per pair: every thread forms conj(HDIFFER * residual) of its receivers
// (vals, [pair][rec][3]), then one lane adds them into the zeroed RHS in
// receiver order, as RESID_DRHS does (repeated DOFs sum in that order;
// conj(a) + conj(b) = conj(a + b) exactly); DOFs outside 1..NDOF are
// skipped and flagged.  est/obs/wghts in the reference's (3, NFREQ, NREC,
// NSRC) layout; ifreq/isrc checked by the host.
__global__ __launch_bounds__(256) void k_cadjlam(int ndof, int nfreq, int nrec, const int *__restrict__ ifreq,
                                                 const int *__restrict__ isrc, const double *__restrict__ freq,
                                                 const double *__restrict__ py, AdjParams prm,
                                                 const double *__restrict__ yrec, const int *__restrict__ mrdof,
                                                 const float2 *__restrict__ est, const float2 *__restrict__ obs,
                                                 const float *__restrict__ wghts, float2 *__restrict__ vals,
                                                 float2 *__restrict__ rhs, int *__restrict__ warn) {
    const int p = blockIdx.x;
    const double f = freq[p];
    const float2 h = hdiffer(prm.ibphase, f);
    float2 *v = vals + (size_t)p * nrec * 3;
    for (int ir = threadIdx.x; ir < nrec; ir += blockDim.x) {
        const size_t base = 3 * ((size_t)(ifreq[p] - 1) + (size_t)nfreq * (ir + (size_t)nrec * (isrc[p] - 1)));
        float2 e[3], o[3], E[3], O[3], R[3];
        for (int i = 0; i < 3; ++i) { e[i] = est[base + i]; o[i] = obs[base + i]; }
        if (prm.lunwrap)
            for (int i = 0; i < 3; ++i) { e[i] = cphm2cm(e[i].x, e[i].y); o[i] = cphm2cm(o[i].x, o[i].y); }
        crotate(prm.az, e[0], e[1], E[0], E[1]);
        E[2] = e[2];
        const float2 cf = cexpi(-(float)(6.2831853071795862 * f * py[p] * yrec[ir]));
        for (int i = 0; i < 3; ++i) { O[i] = cmulf(o[i], cf); E[i] = cmulf(E[i], cf); }
        for (int i = 0; i < 3; ++i)
            R[i] = bpresid(prm.norm, prm.irestp, prm.epss, f, prm.dtmax, wghts[base + i], O[i], E[i]);
        float2 q1, q2;
        crotate(-prm.az, R[0], R[1], q1, q2);
        R[0] = q1; R[1] = q2;
        for (int i = 0; i < 3; ++i) {
            const float2 t = cmulf(h, R[i]);
            v[3 * ir + i] = make_float2(t.x, -t.y);
        }
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    float2 *out = rhs + (size_t)p * ndof;
    for (int k = 0; k < 3 * nrec; ++k) {
        const int d = mrdof[k];
        if (d <= 0 || d > ndof) { atomicOr(warn, 1); continue; }
        out[d - 1].x += v[k].x;
        out[d - 1].y += v[k].y;
    }
}

// BPRESID_COVD4's (objfn.f:635-733) phase gate: the phase difference in
// seconds, formed in single precision
__device__ inline bool covd_gate(float dtmax, double freq, float2 obs, float2 est) {
    const float twopi = 6.283185307179586f, pi = 3.1415926535897931f;
    if (!(dtmax > 0.f)) return false;
    float dphi = atan2r(obs.y, obs.x) - atan2r(est.y, est.x);
    if (dphi < -pi) dphi = dphi + twopi;
    if (dphi > +pi) dphi = dphi - twopi;
    dphi = dphi / (twopi * (float)freq);
    return fabsf(dphi) > dtmax;
}

// LAPLACE_1DV2's (locrec2d.f:334-509) element length between the j-th and
// (j+1)-th observed receivers of a component, whose receiver numbers are
// inv[j] / 3 and inv[j + 1] / 3: the normalised spacings summed in receiver
// order (1 without ldwght)
__device__ inline double covd_h(int ldwght, const double *__restrict__ dx, const int *__restrict__ inv, int j) {
    if (!ldwght) return 1.0;
    double xs = 0.0;
    for (int r = inv[j] / 3, r1 = inv[j + 1] / 3; r < r1; ++r) xs += dx[r];
    return xs;
}

// CADJLAM's right-hand side with the data covariance (adjgrad.f90:219-235,
// 296-300), one workgroup of 256 threads per (ifreq, isrc) pair:
//  1. per receiver: rotation to (N,E,Z), phase shift, CRESID of every
//     component whose shifted observation is non-zero, gated by the phase
//     window (dif, [rec][3]);
//  2. PERM_OBS (locrec2d.f:511-549) of the untransformed observations by a
//     ballot scan per component (taken 3, 2, 1): observed receivers first,
//     then the others; pos[3 irec + i] = position, inv its inverse;
//  3. RESIDV = W C W WORK with WORK = dif at the positions and W the
//     receiver weights there, C the roughening matrix of LAPLACE_1DV2: a
//     chain of two-node elements over each component's observed receivers,
//     numbered consecutively over the components in the same order (so
//     that C's numbering and PERM_OBS's positions coincide only where the
//     earlier components are observed everywhere, as in the reference).
//     C is tridiagonal; each row forms the three products in SGEMV's
//     column order, in single precision without contraction, with C's
//     entries summed in double as the element loop adds them;
//  4. RESID at each receiver's positions, rotated back, HDIFFER, conjugate,
//     and one lane scatters into the zeroed RHS in receiver order.
// Scratch per pair: inv, pos (3 nrec ints), dif, res, vals (3 nrec float2).
__global__ __launch_bounds__(256) void k_cadjlam_covd(
    int ndof, int nfreq, int nrec, const int *__restrict__ ifreq, const int *__restrict__ isrc,
    const double *__restrict__ freq, const double *__restrict__ py, AdjParams prm, int ldwght,
    const double *__restrict__ dx, const double *__restrict__ yrec, const int *__restrict__ mrdof,
    const float2 *__restrict__ est, const float2 *__restrict__ obs, const float *__restrict__ wghts,
    int *__restrict__ sinv, int *__restrict__ spos, float2 *__restrict__ sdif, float2 *__restrict__ sres,
    float2 *__restrict__ vals, float2 *__restrict__ rhs, int *__restrict__ warn) {
    __shared__ int s_na[3], s_wc[4];
    const int p = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int n3 = 3 * nrec;
    const double f = freq[p];
    const size_t gb0 = 3 * (size_t)(ifreq[p] - 1), gstride = 3 * (size_t)nfreq;   // receiver ir at gb0 + ir*gstride
    const size_t gsrc = (size_t)nrec * gstride * (isrc[p] - 1);
    int *inv = sinv + (size_t)p * n3, *pos = spos + (size_t)p * n3;
    float2 *dif = sdif + (size_t)p * n3, *res = sres + (size_t)p * n3, *v = vals + (size_t)p * n3;
    if (tid < 3) s_na[tid] = 0;
    __syncthreads();
    // 1. residuals at the receivers; observed-receiver counts per component
    int cnt[3] = {0, 0, 0};
    for (int ir = tid; ir < nrec; ir += blockDim.x) {
        const size_t base = gsrc + gb0 + ir * gstride;
        float2 e[3], o[3], E[3], O[3];
        for (int i = 0; i < 3; ++i) {
            e[i] = est[base + i];
            o[i] = obs[base + i];
            cnt[i] += cabsf_(o[i]) > 0.f;
        }
        if (prm.lunwrap)
            for (int i = 0; i < 3; ++i) { e[i] = cphm2cm(e[i].x, e[i].y); o[i] = cphm2cm(o[i].x, o[i].y); }
        crotate(prm.az, e[0], e[1], E[0], E[1]);
        E[2] = e[2];
        const float2 cf = cexpi(-(float)(6.2831853071795862 * f * py[p] * yrec[ir]));
        for (int i = 0; i < 3; ++i) {
            O[i] = cmulf(o[i], cf);
            E[i] = cmulf(E[i], cf);
            float2 d = make_float2(0.f, 0.f);
            if (cabsf_(O[i]) > 0.f && !covd_gate(prm.dtmax, f, O[i], E[i])) d = cresid(prm.irestp, O[i], E[i]);
            dif[3 * ir + i] = d;
        }
    }
    for (int i = 0; i < 3; ++i) atomicAdd(&s_na[i], cnt[i]);
    __syncthreads();
    // 2. positions: component c = 2, 1, 0 fills positions b nrec .. (b + 1) nrec - 1, b = 2 - c
    for (int b = 0; b < 3; ++b) {
        const int c = 2 - b, na = s_na[c];
        int run = 0;
        for (int r0 = 0; r0 < nrec; r0 += 256) {
            const int ir = r0 + tid;
            const bool act = ir < nrec && cabsf_(obs[gsrc + gb0 + ir * gstride + c]) > 0.f;
            const unsigned long long m = __ballot(act);
            if (lane == 0) s_wc[wv] = __popcll(m);
            __syncthreads();
            int off = run;
            for (int j = 0; j < wv; ++j) off += s_wc[j];
            run += s_wc[0] + s_wc[1] + s_wc[2] + s_wc[3];
            __syncthreads();
            if (ir < nrec) {
                const int aex = off + __popcll(m & ((1ull << lane) - 1));   // observed receivers before ir
                const int l = b * nrec + (act ? aex : na + (ir - aex));
                pos[3 * ir + c] = l;
                inv[l] = 3 * ir + c;
            }
        }
    }
    __syncthreads();
    // 3. RESIDV = W C W WORK on the tridiagonal C
    const int na0 = s_na[2], na1 = s_na[1], nall = s_na[2] + s_na[1] + s_na[0];
    for (int l = tid; l < n3; l += blockDim.x) {
        float2 acc = make_float2(0.f, 0.f);
        if (l < nall) {
            const int b = l < na0 ? 0 : (l < na0 + na1 ? 1 : 2);
            const int kk = l - (b == 0 ? 0 : (b == 1 ? na0 : na0 + na1)), na = s_na[2 - b];
            const int *cinv = inv + b * nrec;       // observed receivers of the component, in order
            double cm = 0.0, cd = 0.0, cp = 0.0;    // C(l, l-1), C(l, l), C(l, l+1)
            if (kk > 0) {
                const double t = 2.0 / covd_h(ldwght, dx, cinv, kk - 1);
                cm = t * -0.5;
                cd = cd + t * 0.5;
            }
            if (kk + 1 < na) {
                const double t = 2.0 / covd_h(ldwght, dx, cinv, kk);
                cd = cd + t * 0.5;
                cp = t * -0.5;
            }
            const float wl = wghts[gsrc + gb0 + (inv[l] / 3) * gstride + inv[l] % 3];
            for (int j = -1; j <= 1; ++j) {
                const int k = l + j;
                if (k < 0 || k >= n3) continue;
                const double cv = j < 0 ? cm : (j == 0 ? cd : cp);
                if (cv == 0.0) continue;
                const float wk = wghts[gsrc + gb0 + (inv[k] / 3) * gstride + inv[k] % 3];
                const float a = __fmul_rn(__fmul_rn(wl, (float)cv), wk);
                const float2 x = dif[inv[k]];
                acc.x = __fadd_rn(acc.x, __fmul_rn(x.x, a));
                acc.y = __fadd_rn(acc.y, __fmul_rn(x.y, a));
            }
        }
        res[l] = acc;
    }
    __syncthreads();
    // 4. back to the receivers, rotated back, HDIFFER, conjugate
    const float2 h = hdiffer(prm.ibphase, f);
    for (int ir = tid; ir < nrec; ir += blockDim.x) {
        float2 R[3];
        for (int i = 0; i < 3; ++i) R[i] = res[pos[3 * ir + i]];
        float2 q1, q2;
        crotate(-prm.az, R[0], R[1], q1, q2);
        R[0] = q1; R[1] = q2;
        for (int i = 0; i < 3; ++i) {
            const float2 t = cmulf(h, R[i]);
            v[3 * ir + i] = make_float2(t.x, -t.y);
        }
    }
    __syncthreads();
    if (tid != 0) return;
    float2 *out = rhs + (size_t)p * ndof;
    for (int k = 0; k < n3; ++k) {
        const int d = mrdof[k];
        if (d <= 0 || d > ndof) { atomicOr(warn, 1); continue; }
        out[d - 1].x += v[k].x;
        out[d - 1].y += v[k].y;
    }
}

// ---------------------------------------------------------------------------
// Bielak plane-wave sources for a batch.  FILL_UWAVE_QMESH
// (disp25.f90:710-908): u[d][w] = pol_w(comp d) exp(-i (kx_w x_d + kz_w z_d))
// with (x_d, z_d) the DOF's node as the first element holding it places it
// (the host's table); the phase products rounded as the reference rounds
// them (no contraction).  pw[w] = (kx, kz, pu, pv, pw): polarisation
// (cos baz sin aoi, sin baz sin aoi, cos aoi).  [dof][wave] layout.
__global__ __launch_bounds__(256) void k_uwave(int64_t ndof, int nw, const double2 *__restrict__ xz,
                                               const signed char *__restrict__ comp,
                                               const double *__restrict__ pw, double2 *__restrict__ ue) {
    const int64_t tot = ndof * nw;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < tot; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t d = i / nw;
        const int w = (int)(i - d * nw);
        const double *p = pw + 5 * (size_t)w;
        const double2 c = xz[d];
        const double t = __dsub_rn(__dmul_rn(-p[0], c.x), __dmul_rn(p[1], c.y));
        double sn, cs;
        sincos(t, &sn, &cs);
        const double a = p[2 + comp[d]];
        ue[i] = make_double2(__dmul_rn(a, cs), __dmul_rn(a, sn));
    }
}

// rows[k] (0-based DOFs) of a [dof][pair] vector negated
__global__ __launch_bounds__(256) void k_negate_rows(int nrow, int npair, const int *__restrict__ rows,
                                                     double2 *__restrict__ y) {
    const int64_t tot = (int64_t)nrow * npair;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < tot; i += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / npair), p = (int)(i - (int64_t)r * npair);
        double2 *v = y + (size_t)rows[r] * npair + p;
        *v = make_double2(-v->x, -v->y);
    }
}

// ---------------------------------------------------------------------------
// small utilities
template <typename T>
__global__ void k_fill(T *p, size_t n, T v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// per-pair max over all DOFs of |Re x|, |Im x| for x[dof][pair]; non-negative
// doubles order like their bit patterns, so an integer max reduces them
__global__ void k_colmax(int64_t ndof, int npair, const double2 *__restrict__ x,
                         unsigned long long *__restrict__ cm) {
    const int pair = blockIdx.x * 64 + (threadIdx.x & 63);
    if (pair >= npair) return;
    const int wpb = blockDim.x / 64;
    double mx = 0.0;
    for (int64_t d = (int64_t)blockIdx.y * wpb + threadIdx.x / 64; d < ndof; d += (int64_t)gridDim.y * wpb) {
        const double2 v = x[d * npair + pair];
        mx = fmax(mx, fmax(fabs(v.x), fabs(v.y)));
    }
    atomicMax(cm + pair, (unsigned long long)__double_as_longlong(mx));
}

// [pair][dof] <-> [dof][pair] transposes for the host convenience wrappers
// tiles numbered along a 1-D grid (either dimension may be long)
template <typename T>
__global__ void k_transpose_c(const T *__restrict__ in, T *__restrict__ out, int rows, int cols) {
    __shared__ T tile[16][17];
    const int ntc = (cols + 15) / 16;
    const int tr = blockIdx.x / ntc, tcol = blockIdx.x % ntc;
    int c = tcol * 16 + threadIdx.x, r = tr * 16 + threadIdx.y;
    if (r < rows && c < cols) tile[threadIdx.y][threadIdx.x] = in[(size_t)r * cols + c];
    __syncthreads();
    c = tr * 16 + threadIdx.x;
    r = tcol * 16 + threadIdx.y;
    if (r < cols && c < rows) out[(size_t)r * rows + c] = tile[threadIdx.x][threadIdx.y];
}

// MFMA lane-layout self check: D = A(16x4) B(4x16) for both data types
__global__ void k_mfma_probe(double *outd, float *outf) {
    const int l = threadIdx.x;
    const int row = l & 15, k = l >> 4;
    double a = 1.0 + row + 16.0 * k, b = 1.0 + 0.5 * (l & 15) - 3.0 * k;
    MfmaTraits<double>::v4 cd = {0, 0, 0, 0};
    cd = MfmaTraits<double>::mma(a, b, cd);
    MfmaTraits<float>::v4 cf = {0, 0, 0, 0};
    cf = MfmaTraits<float>::mma((float)a, (float)b, cf);
    for (int i = 0; i < 4; ++i) {
        outd[l * 4 + i] = cd[i];
        outf[l * 4 + i] = cf[i];
    }
}

// Workgroup residency probe: every wave spins a bounded number of shader
// cycles, then lane 0 records the workgroup's realtime start and end and
// its hardware location (HW_ID: CU, shader array, engine; XCC_ID) as
// rec[3 wg + 0..2] (wave 0 only).  The register count is forced by an
// inline-assembly clobber of the highest register (VG - 1), so the kernel
// holds exactly the VGPR budget of the kernels it stands for.
#define F25_RESID_PROBE(VG, VGM1)                                                                   \
    template <int NT>                                                                               \
    __global__ __launch_bounds__(NT) void k_resid_probe_##VG(unsigned long long *rec, int nwg,     \
                                                             unsigned long long spin) {             \
        const unsigned long long r0 = __builtin_amdgcn_s_memrealtime();                             \
        const unsigned long long t0 = __builtin_amdgcn_s_memtime();                                 \
        unsigned long long t = t0;                                                                  \
        for (int it = 0; t - t0 < spin && it < (1 << 22); ++it) {                                   \
            __builtin_amdgcn_s_sleep(1);                                                            \
            t = __builtin_amdgcn_s_memtime();                                                       \
        }                                                                                           \
        asm volatile("; holds the register budget" ::: #VGM1);                                      \
        const unsigned long long r1 = __builtin_amdgcn_s_memrealtime();                             \
        if (threadIdx.x == 0 && (int)blockIdx.x < nwg) {                                            \
            unsigned long long *o = rec + 3 * (size_t)blockIdx.x;                                   \
            o[0] = r0;                                                                              \
            o[1] = r1;                                                                              \
            o[2] = (unsigned long long)__builtin_amdgcn_s_getreg((31 << 11) | 4) |                  \
                   ((unsigned long long)__builtin_amdgcn_s_getreg((31 << 11) | 20) << 32);          \
        }                                                                                           \
    }
F25_RESID_PROBE(128, v127)
F25_RESID_PROBE(168, v167)
#undef F25_RESID_PROBE

}  // namespace f25k
