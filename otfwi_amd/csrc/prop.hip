// Fused time-step kernels of the split-field PML acoustic propagator.
//
// Scheme (reference: simulator/src/regression_PML_single.cu:447-615 with the
// prec_deriv==4 coefficients of simulator/src/deriv_macros.h:84-108):
//   P half-step  Px(t) = Px(t-1)*kx + m*D-x Ax(t-1),  Pz(t) = Pz(t-1)*kz + m*D-z Az(t-1)
//   A half-step  Ax(t) = Ax(t-1)*kx + cA*D+x P(t),    Az(t) = Az(t-1)*kz + cA*D+z P(t)
// with P = Px+Pz, m = (c^2 dt)/dx, cA = dt/dx, k = 1 - d*dt looked up one
// cell further along its axis, and every cell outside the box
// 4 <= x < X-4, 4 <= z < Z-4 held at 0.
//
// Storage: P = Px+Pz is kept on every cell; the split component Px (the
// boundary memory variable "psi") is non-zero only on damped cells, where
// kx != 1 or kz != 1, and only strip tiles touch it.  Fields are padded
// planes (otfwi_internal.h) whose padding and dead ring are always 0, so all
// loads are unconditional, 16-byte vectors.
//
// One launch advances every shot of a batch by one full time step.  A
// 256-thread work-group owns a 64 x TH tile:
//   1. it loads A(t-1) with an 8-cell halo into LDS and P(t-1), m for the
//      tile plus a 4-cell halo (the "E box") into registers -- every global
//      load of the tile is issued before the first barrier;
//   2. E phase: P(t) on the E box, in column quads with a register window
//      along z, into LDS (source injection, receiver sampling and the
//      boundary-ring save happen here, on the cells where they occur);
//   3. owned phase: A(t) on the tile.
// The per-step HBM traffic is the algorithmic minimum (read P, Ax, Az, m;
// write P, Ax, Az) plus halo re-reads that hit in L2.
//
// The two-step kernels (pair2_kernel, at the end of this file) advance two
// time steps per launch with z-marching work-groups that keep the
// intermediate state on chip, for the forward field, the adjoint field and
// the reconstructed forward field; they halve the HBM traffic per step and
// are used for launches large enough to fill the GPU.
#include "otfwi_internal.h"

namespace otfwi {

template <typename R> struct Vec4T;
template <> struct Vec4T<float> { using T = float4; };
template <> struct Vec4T<double> { using T = double4; };
template <typename R> using V4 = typename Vec4T<R>::T;

template <typename R>
__device__ __forceinline__ V4<R> ld4(const R *p) { return *reinterpret_cast<const V4<R> *>(p); }
template <typename R>
__device__ __forceinline__ void st4(R *p, const V4<R> &v) { *reinterpret_cast<V4<R> *>(p) = v; }
// Global store of a step's output with the non-temporal hint: the outputs of
// one launch (GBs) are read back only by the next launch, long after they
// would have left L2.  Measured on MI355X: +6 % on a plain 4-read/3-write
// streaming kernel, neutral on the step kernels (same launch times as plain
// stores).
template <typename R> struct ExtV4;
template <> struct ExtV4<float> { typedef float T __attribute__((ext_vector_type(4))); };
template <> struct ExtV4<double> { typedef double T __attribute__((ext_vector_type(4))); };
template <typename R>
__device__ __forceinline__ void stg4(R *p, const V4<R> &v)
{
    const typename ExtV4<R>::T e = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(e, reinterpret_cast<typename ExtV4<R>::T *>(p));
}
template <typename R>
__device__ __forceinline__ void unpack(const V4<R> &v, R *a) { a[0] = v.x; a[1] = v.y; a[2] = v.z; a[3] = v.w; }
template <typename R>
__device__ __forceinline__ V4<R> pack(const R *a) { V4<R> v; v.x = a[0]; v.y = a[1]; v.z = a[2]; v.w = a[3]; return v; }

// Correctly rounded product that the compiler may not fuse into a following
// add.  Every update below is written as fma(coef, derivative, mul_rn(old,
// k)): when the damping factor k is the constant 1 (interior tiles) and when
// it is loaded (strip tiles) the same roundings happen, so results do not
// depend on which tiles are strip tiles, i.e. on the tile height.
__device__ __forceinline__ float mul_rn(float a, float b) { return __fmul_rn(a, b); }
__device__ __forceinline__ double mul_rn(double a, double b) { return __dmul_rn(a, b); }
__device__ __forceinline__ float add_rn(float a, float b) { return __fadd_rn(a, b); }
__device__ __forceinline__ double add_rn(double a, double b) { return __dadd_rn(a, b); }

// Coefficients of the 8th-order staggered first derivative exactly as the
// reference writes them (double literals, deriv_macros.h:87-94).
constexpr double K1 = 1.196289e+00, K2 = -7.975260e-02, K3 = 9.570313e-03, K4 = -6.975446e-04;

// sum_k c_k (v[3+k] - v[4-k]) for taps v[j] = f(x0 + j): the backward
// operator D- at x = x0+4 and the forward operator D+ at x = x0+3.
//
// FP32: the coefficients are the FP32 roundings of the reference's literals
// plus the low part of c1 folded in first, and the terms are accumulated
// smallest first.  Rounding the coefficients to FP32 alone changes the stencil
// systematically (the C1 adjoint gradient moves by ~1.7e-4 relative); the
// c1 low part restores it to the reference's own FP32 level (~2e-5 from FP64);
// the low parts of c2..c4 lie below FP32 resolution of the derivative.
template <typename R>
__device__ __forceinline__ R stag(R v0, R v1, R v2, R v3, R v4, R v5, R v6, R v7);

template <>
__device__ __forceinline__ float stag<float>(float v0, float v1, float v2, float v3, float v4,
                                             float v5, float v6, float v7)
{
    constexpr float h1 = (float)K1, h2 = (float)K2, h3 = (float)K3, h4 = (float)K4;
    constexpr float l1 = (float)(K1 - (double)h1);
    const float d1 = v4 - v3, d2 = v5 - v2, d3 = v6 - v1, d4 = v7 - v0;
    float s = l1 * d1;
    s = fmaf(h4, d4, s);
    s = fmaf(h3, d3, s);
    s = fmaf(h2, d2, s);
    s = fmaf(h1, d1, s);
    return s;
}

template <>
__device__ __forceinline__ double stag<double>(double v0, double v1, double v2, double v3,
                                               double v4, double v5, double v6, double v7)
{
    double s = K4 * (v7 - v0);
    s = fma(K3, v6 - v1, s);
    s = fma(K2, v5 - v2, s);
    s = fma(K1, v4 - v3, s);
    return s;
}

// 9-point 8th-order second difference numerator (deriv_macros.h:110-135),
// to be divided by 5040.
template <typename R>
__device__ __forceinline__ R lap9(R m4, R m3, R m2, R m1, R c, R p1, R p2, R p3, R p4)
{
    R s = R(-9) * (m4 + p4);
    s = fma(R(128), m3 + p3, s);
    s = fma(R(-1008), m2 + p2, s);
    s = fma(R(8064), m1 + p1, s);
    s = fma(R(-14350), c, s);
    return s;
}

// Work-groups are dispatched round-robin over the 8 XCDs (block b runs on
// XCD b % 8).  Give each XCD a contiguous range of logical tiles so that
// spatial neighbours, whose halos overlap, share one L2.
__device__ __forceinline__ int xcd_remap(int b, int nb)
{
    const int q = nb >> 3, r = nb & 7;
    const int xcd = b & 7, k = b >> 3;
    const int base = (xcd < r) ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q;
    return base + k;
}

// Staging of a (ROWS x 4*NQ) block of a padded plane into LDS through
// registers: every thread issues all of its loads before any store.
template <typename R, int ROWS, int NQ>
struct Stage {
    static constexpr int N = ROWS * NQ;
    static constexpr int K = (N + NTHREADS - 1) / NTHREADS;
    V4<R> r[K];
    // base: the padded plane; (z_first, x_first) in cell coordinates
    __device__ __forceinline__ void load(const R *base, int ld, int z_first, int x_first, int tid)
    {
        const unsigned o0 = (unsigned)((z_first + ZOFF) * ld + x_first + XOFF);
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int i = min(tid + k * NTHREADS, N - 1);
            const int row = i / NQ, q = i - row * NQ;
            r[k] = ld4(base + (o0 + (unsigned)(row * ld + 4 * q)));
        }
    }
    template <int W>
    __device__ __forceinline__ void store(R (*lds)[W], int tid) const
    {
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int i = tid + k * NTHREADS;
            if (i < N) {
                const int row = i / NQ, q = i - row * NQ;
                st4(&lds[row][4 * q], r[k]);
            }
        }
    }
};

// d[i] = stag(v[o+i .. o+i+7]), i = 0..3, from a 12-value row segment.
template <typename R, int O>
__device__ __forceinline__ void stag_row(const R (&v)[12], R (&d)[4])
{
#pragma unroll
    for (int i = 0; i < 4; i++)
        d[i] = stag<R>(v[O + i], v[O + i + 1], v[O + i + 2], v[O + i + 3], v[O + i + 4],
                       v[O + i + 5], v[O + i + 6], v[O + i + 7]);
}

template <typename R>
__device__ __forceinline__ void row12(const R *row, R (&v)[12])
{
    unpack<R>(ld4(row), v);
    unpack<R>(ld4(row + 4), v + 4);
    unpack<R>(ld4(row + 8), v + 8);
}

// d[i] = stag over a z window of 8 quads (column i of each).
template <typename R>
__device__ __forceinline__ void stag_col(const V4<R> (&w)[8], R (&d)[4])
{
    R a[8][4];
#pragma unroll
    for (int j = 0; j < 8; j++) unpack<R>(w[j], a[j]);
#pragma unroll
    for (int i = 0; i < 4; i++)
        d[i] = stag<R>(a[0][i], a[1][i], a[2][i], a[3][i], a[4][i], a[5][i], a[6][i], a[7][i]);
}

// Work-group barrier ordering LDS only.  The step kernels exchange data
// through LDS alone, so their barriers need not wait for outstanding global
// loads (the next tile's prefetch) or stores (results already computed).
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ bool in_box(int x, int z, int X, int Z)
{
    return x >= BOX && z >= BOX && x < X - BOX && z < Z - BOX;
}

// Row assignment of the E phase: the EH rows of the 16 central column quads
// are split over 16 row groups as evenly as possible; the 2*TH side cells
// (columns x0-4..x0-1 and x0+64..x0+67 on the tile rows) go one quad-row
// each to the last 2*TH threads, whose groups hold the shorter runs.
template <int TH>
struct EMap {
    static constexpr int EH = TH + 8;
    static constexpr int RB = EH / 16, RX = EH % 16;
    static constexpr int RMAX = RB + (RX ? 1 : 0);
    static constexpr int NSIDE = 2 * TH;
    static_assert(NSIDE <= NTHREADS, "side cells exceed the work-group");
    static __device__ __forceinline__ int start(int g) { return g * RB + min(g, RX); }
    static __device__ __forceinline__ int count(int g) { return RB + (g < RX ? 1 : 0); }
};

// ---------------------------------------------------------------------------
// forward step
// ---------------------------------------------------------------------------

template <typename R>
struct FwdCtx {
    const FwdArgs<R> &a;
    int s, X, Z, ld;
    uint8_t fl;
    const R *psi0;
    R *psi1, *P1;
    const int *evt;   // event map at the plane origin (cell (0,0))
};

// P(t) for one E quad (cells x..x+3 on row z), given D-x Ax and D-z Az.
// Returns the four values (0 outside the updated box) and performs the
// source injection, receiver sampling, ring save and psi update there.
template <typename R>
__device__ __forceinline__ void p_quad(const FwdCtx<R> &c, int x, int z, bool owned,
                                       const R (&dAx)[4], const R (&dAz)[4], const V4<R> &pp4,
                                       const V4<R> &mm4, R (&p)[4])
{
    const FwdArgs<R> &a = c.a;
    R pp[4], mm[4], kx[4] = {R(1), R(1), R(1), R(1)}, psi[4] = {R(0), R(0), R(0), R(0)};
    unpack<R>(pp4, pp);
    unpack<R>(mm4, mm);
    R kz = R(1);
    const unsigned gi = (unsigned)((z + ZOFF) * c.ld + x + XOFF);
    if (c.fl & TF_STRIP) {
        unpack<R>(ld4(a.kx + (unsigned)(XOFF + x)), kx);
        kz = a.kz[ZOFF + z];
        unpack<R>(ld4(c.psi0 + gi), psi);
    }
    R nx[4], nz[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        nx[i] = fma(mm[i], dAx[i], mul_rn(psi[i], kx[i]));
        nz[i] = fma(mm[i], dAz[i], mul_rn(pp[i] - psi[i], kz));
    }
    if (c.fl & TF_EVENT) {
        int ev[4];
        {
            const int4 e4 = *reinterpret_cast<const int4 *>(c.evt + gi);
            ev[0] = e4.x; ev[1] = e4.y; ev[2] = e4.z; ev[3] = e4.w;
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            if (ev[i] < 0) continue;
            const PointEvt pe = a.pts[ev[i]];
            if (pe.src >= 0) {
                const int ss = a.shot_src[c.s];
                if (ss < 0 || ss == pe.src) {
                    const R h = a.icum[(size_t)pe.src * a.T + a.t] * R(0.5);
                    nx[i] = add_rn(nx[i], h);
                    nz[i] = add_rn(nz[i], h);
                }
            }
            if (owned) {
                const R pv = add_rn(nx[i], nz[i]);
                if (pe.rx >= 0 && a.rec) a.rec[((size_t)c.s * a.n_rx + pe.rx) * a.T + a.t] = pv;
                if (pe.ring >= 0 && a.ring)
                    a.ring[((size_t)c.s * a.T + a.t) * a.n_ring + pe.ring] = pv;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 4; i++) p[i] = add_rn(nx[i], nz[i]);
    if (c.fl & TF_EDGE) {
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (!in_box(x + i, z, c.X, c.Z)) { p[i] = R(0); nx[i] = R(0); }
    }
    if (owned) {
        stg4(c.P1 + gi, pack<R>(p));
        if (c.fl & TF_STRIP) {
            R ps[4];
#pragma unroll
            for (int i = 0; i < 4; i++) ps[i] = (kx[i] != R(1) || kz != R(1)) ? nx[i] : R(0);
            stg4(c.psi1 + gi, pack<R>(ps));
        }
    }
}

// Per-thread register image of everything one tile loads from HBM, so the
// loads of the next tile can be in flight while the current one computes.
template <typename R, int TH>
struct FwdLoads {
    using M = EMap<TH>;
    static constexpr int EH = M::EH, AZH = TH + 16;
    Stage<R, EH, AXW / 4> stx;
    Stage<R, AZH, EW / 4> stz;
    V4<R> pp[M::RMAX], mm[M::RMAX], spp, smm;

    // sb: element offset of the shot's planes
    __device__ __forceinline__ void load(const FwdArgs<R> &a, int x0, int z0, size_t sb, int tid)
    {
        const int ld = a.ld, q = tid & 15, g = tid >> 4;
        stx.load(a.Ax0 + sb, ld, z0 - 4, x0 - 8, tid);
        stz.load(a.Az0 + sb, ld, z0 - 8, x0 - 4, tid);
        const R *P0 = a.P0 + sb, *cm = a.cm;
        const int er0 = M::start(g);
#pragma unroll
        for (int j = 0; j < M::RMAX; j++) {
            const int er = min(er0 + j, EH - 1);
            const unsigned gi = (unsigned)((z0 - 4 + er + ZOFF) * ld + x0 + 4 * q + XOFF);
            pp[j] = ld4(P0 + gi);
            mm[j] = ld4(cm + gi);
        }
        const int sidx = tid - (NTHREADS - M::NSIDE);
        const int s_er = 4 + (sidx >= 0 ? sidx % TH : 0);
        const int s_q = (sidx >= TH) ? 17 : 0;
        const unsigned gi = (unsigned)((z0 - 4 + s_er + ZOFF) * ld + x0 - 4 + 4 * s_q + XOFF);
        spp = ld4(P0 + gi);
        smm = ld4(cm + gi);
    }
};

// Tile (s, tl) from the logical tile index L.
struct TileId {
    int s, tl, x0, z0;
    size_t sb;   // element offset of the shot's planes
};
template <int TH>
__device__ __forceinline__ TileId tile_id(int L, int ntx, int ntz, int ld, long long sstride)
{
    TileId t;
    const int tiles = ntx * ntz;
    t.s = L / tiles;
    t.tl = L - t.s * tiles;
    const int tz = t.tl / ntx, tx = t.tl - tz * ntx;
    t.x0 = tx * TW;
    t.z0 = tz * TH;
    t.sb = (size_t)t.s * (size_t)sstride;
    (void)ld;
    return t;
}

// E phase + owned phase of one tile whose A(t-1) is staged in LDS and whose
// P(t-1), m quads are in `ld`.  Ends after the owned-phase stores (callers
// that reuse the LDS must synchronize first).
template <typename R, int TH>
__device__ __forceinline__ void fwd_compute(const FwdArgs<R> &a, const TileId &T,
                                            const FwdLoads<R, TH> &L, int tid, R (*sAx)[AXW],
                                            R (*sAz)[EW], R (*sP)[EW])
{
    using M = EMap<TH>;
    constexpr int AZH = TH + 16, RO = TH / 16;
    const int ld = a.ld, q = tid & 15, g = tid >> 4;
    const int x0 = T.x0, z0 = T.z0;
    FwdCtx<R> c{a, T.s, a.X, a.Z, ld, a.tflags[T.tl], a.Px0 + T.sb, a.Px1 + T.sb, a.P1 + T.sb,
                a.evt};
    const int er0 = M::start(g), ern = M::count(g);
    const int sidx = tid - (NTHREADS - M::NSIDE);
    const bool has_side = sidx >= 0;
    const int s_er = 4 + (has_side ? sidx % TH : 0);
    const int s_q = (has_side && sidx >= TH) ? 17 : 0;

    // ---- E phase: P(t) ------------------------------------------------
    {
        const int Q = q + 1;   // central E quads 1..16
        V4<R> w[8];
#pragma unroll
        for (int j = 0; j < 8; j++) w[j] = ld4(&sAz[min(er0 + j, AZH - 1)][4 * Q]);
#pragma unroll
        for (int j = 0; j < M::RMAX; j++) {
            if (j < ern) {
                const int er = er0 + j;
                R v[12], dAx[4], dAz[4], p[4];
                row12<R>(&sAx[er][4 * Q], v);
                stag_row<R, 0>(v, dAx);
                stag_col<R>(w, dAz);
                p_quad<R>(c, x0 + 4 * q, z0 - 4 + er, er >= 4 && er < TH + 4, dAx, dAz, L.pp[j],
                          L.mm[j], p);
                st4(&sP[er][4 * Q], pack<R>(p));
            }
            if (j + 1 < M::RMAX) {
#pragma unroll
                for (int k = 0; k < 7; k++) w[k] = w[k + 1];
                w[7] = ld4(&sAz[min(er0 + j + 8, AZH - 1)][4 * Q]);
            }
        }
        if (has_side) {
            R v[12], dAx[4], dAz[4], p[4];
            row12<R>(&sAx[s_er][4 * s_q], v);
            stag_row<R, 0>(v, dAx);
            V4<R> ws[8];
#pragma unroll
            for (int j = 0; j < 8; j++) ws[j] = ld4(&sAz[s_er + j][4 * s_q]);
            stag_col<R>(ws, dAz);
            p_quad<R>(c, x0 - 4 + 4 * s_q, z0 - 4 + s_er, false, dAx, dAz, L.spp, L.smm, p);
            st4(&sP[s_er][4 * s_q], pack<R>(p));
        }
    }
    lds_barrier();

    // ---- owned phase: A(t) ---------------------------------------------
    {
        const int x = x0 + 4 * q;
        const int or0 = g * RO;
        R kx[4] = {R(1), R(1), R(1), R(1)};
        if (c.fl & TF_STRIP) unpack<R>(ld4(a.kx + (unsigned)(XOFF + x)), kx);
        V4<R> w[8];
#pragma unroll
        for (int j = 0; j < 8; j++) w[j] = ld4(&sP[or0 + 1 + j][4 * q + 4]);
        R *Ax1 = a.Ax1 + T.sb, *Az1 = a.Az1 + T.sb;
#pragma unroll
        for (int j = 0; j < RO; j++) {
            const int orr = or0 + j, z = z0 + orr;
            R v[12], dPx[4], dPz[4], ax[4], az[4], nax[4], naz[4];
            row12<R>(&sP[orr + 4][4 * q], v);
            stag_row<R, 1>(v, dPx);
            stag_col<R>(w, dPz);
            if (j + 1 < RO) {
#pragma unroll
                for (int k = 0; k < 7; k++) w[k] = w[k + 1];
                w[7] = ld4(&sP[orr + 9][4 * q + 4]);
            }
            unpack<R>(ld4(&sAx[orr + 4][4 * q + 8]), ax);
            unpack<R>(ld4(&sAz[orr + 8][4 * q + 4]), az);
            const R kz = (c.fl & TF_STRIP) ? a.kz[ZOFF + z] : R(1);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                nax[i] = fma(a.cA, dPx[i], mul_rn(ax[i], kx[i]));
                naz[i] = fma(a.cA, dPz[i], mul_rn(az[i], kz));
            }
            if (c.fl & TF_EDGE) {
#pragma unroll
                for (int i = 0; i < 4; i++)
                    if (!in_box(x + i, z, c.X, c.Z)) { nax[i] = R(0); naz[i] = R(0); }
            }
            const unsigned gi = (unsigned)((z + ZOFF) * ld + x + XOFF);
            stg4(Ax1 + gi, pack<R>(nax));
            stg4(Az1 + gi, pack<R>(naz));
        }
    }
}

#define FWD_LDS(TH)                                                   \
    __shared__ __align__(16) R sAx[EMap<TH>::EH][AXW];                \
    __shared__ __align__(16) R sAz[TH + 16][EW];                      \
    __shared__ __align__(16) R sP[EMap<TH>::EH][EW]

// One tile per work-group.
template <typename R, int TH>
__global__ void __launch_bounds__(NTHREADS)
fwd_step_kernel(const FwdArgs<R> a)
{
    static_assert(TH % 16 == 0, "tile height must be a multiple of 16");
    FWD_LDS(TH);
    const int tid = threadIdx.x;
    const TileId T = tile_id<TH>(xcd_remap(blockIdx.x, gridDim.x), a.ntx, a.ntz, a.ld, a.sstride);
    FwdLoads<R, TH> L;
    L.load(a, T.x0, T.z0, T.sb, tid);
    L.stx.store(sAx, tid);
    L.stz.store(sAz, tid);
    lds_barrier();
    fwd_compute<R, TH>(a, T, L, tid, sAx, sAz, sP);
}

// Range of logical tiles of XCD `xcd` when nb tiles are split over 8 XCDs in
// contiguous chunks (the same split xcd_remap makes).
__device__ __forceinline__ void xcd_chunk(int xcd, int nb, int &lo, int &hi)
{
    const int q = nb >> 3, r = nb & 7;
    lo = (xcd < r) ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q;
    hi = lo + q + (xcd < r ? 1 : 0);
}

// Persistent work-groups: a grid of a few work-groups per CU walks the tiles
// of its XCD's chunk with stride (work-groups on that XCD), so the tiles in
// flight at one time are neighbours sharing halos in L2.  The loads of the
// next tile are issued right after the current tile is in LDS and overlap
// its computation.
template <typename R, int TH>
__global__ void __launch_bounds__(NTHREADS)
fwd_step_persistent(const FwdArgs<R> a)
{
    FWD_LDS(TH);
    const int tid = threadIdx.x;
    const int nb = a.ntx * a.ntz * a.nshots;
    const int xcd = blockIdx.x & 7, k = blockIdx.x >> 3;
    const int nk = (gridDim.x - xcd + 7) >> 3;   // work-groups on this XCD
    int lo, hi;
    xcd_chunk(xcd, nb, lo, hi);
    int Lt = lo + k;
    FwdLoads<R, TH> L;
    TileId T = tile_id<TH>(min(Lt, nb - 1), a.ntx, a.ntz, a.ld, a.sstride);
    if (Lt < hi) L.load(a, T.x0, T.z0, T.sb, tid);
    for (; Lt < hi; Lt += nk) {
        L.stx.store(sAx, tid);
        L.stz.store(sAz, tid);
        FwdLoads<R, TH> C;   // this tile's register part
#pragma unroll
        for (int j = 0; j < EMap<TH>::RMAX; j++) { C.pp[j] = L.pp[j]; C.mm[j] = L.mm[j]; }
        C.spp = L.spp;
        C.smm = L.smm;
        const TileId Tc = T;
        lds_barrier();
        if (Lt + nk < hi) {
            T = tile_id<TH>(Lt + nk, a.ntx, a.ntz, a.ld, a.sstride);
            L.load(a, T.x0, T.z0, T.sb, tid);
        }
        fwd_compute<R, TH>(a, Tc, C, tid, sAx, sAz, sP);
        lds_barrier();
    }
}

// ---------------------------------------------------------------------------
// adjoint step
// ---------------------------------------------------------------------------
// Reference phase 2 (regression_PML_single.cu:793-823): the adjoint field
// lambda advances with the damped forward operator and takes the running sum
// of the time-flipped adjoint source at the receivers; the forward field is
// reconstructed backwards with the undamped reverse operator and the saved
// boundary ring; grad += lambda_P(t) * lap8(U(t)) on the tile.

template <typename R>
struct AdjCtx {
    const AdjArgs<R> &a;
    int s, X, Z, ld, trev;
    uint8_t fl;
    const R *psi0;
    R *psi1, *LP1, *U1;
    const int *evt;   // event map at the plane origin (cell (0,0))
};

template <typename R>
__device__ __forceinline__ void lu_quad(const AdjCtx<R> &c, int x, int z, bool owned,
                                        const R (&dLx)[4], const R (&dLz)[4], const R (&dUx)[4],
                                        const R (&dUz)[4], const V4<R> &lp4, const V4<R> &up4,
                                        const V4<R> &mm4, R (&lp)[4], R (&up)[4])
{
    const AdjArgs<R> &a = c.a;
    R lpp[4], upp[4], mm[4], kx[4] = {R(1), R(1), R(1), R(1)}, psi[4] = {R(0), R(0), R(0), R(0)};
    unpack<R>(lp4, lpp);
    unpack<R>(up4, upp);
    unpack<R>(mm4, mm);
    R kz = R(1);
    const unsigned gi = (unsigned)((z + ZOFF) * c.ld + x + XOFF);
    if (c.fl & TF_STRIP) {
        unpack<R>(ld4(a.kx + (unsigned)(XOFF + x)), kx);
        kz = a.kz[ZOFF + z];
        unpack<R>(ld4(c.psi0 + gi), psi);
    }
    R nx[4], nz[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        nx[i] = fma(mm[i], dLx[i], mul_rn(psi[i], kx[i]));
        nz[i] = fma(mm[i], dLz[i], mul_rn(lpp[i] - psi[i], kz));
        up[i] = fma(-mm[i], dUx[i] + dUz[i], upp[i]);
    }
    if (c.fl & TF_EVENT) {
        int ev[4];
        {
            const int4 e4 = *reinterpret_cast<const int4 *>(c.evt + gi);
            ev[0] = e4.x; ev[1] = e4.y; ev[2] = e4.z; ev[3] = e4.w;
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            if (ev[i] < 0) continue;
            const PointEvt pe = a.pts[ev[i]];
            if (pe.rx >= 0) {
                const R h = a.jcum[((size_t)c.s * a.n_rx + pe.rx) * a.T + a.t] * R(0.5);
                nx[i] = add_rn(nx[i], h);
                nz[i] = add_rn(nz[i], h);
            }
            if (pe.ring >= 0) up[i] = a.ring[((size_t)c.s * a.T + c.trev) * a.n_ring + pe.ring];
        }
    }
#pragma unroll
    for (int i = 0; i < 4; i++) lp[i] = add_rn(nx[i], nz[i]);
    if (c.fl & TF_EDGE) {
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (!in_box(x + i, z, c.X, c.Z)) { lp[i] = R(0); up[i] = R(0); nx[i] = R(0); }
    }
    if (owned) {
        stg4(c.LP1 + gi, pack<R>(lp));
        stg4(c.U1 + gi, pack<R>(up));
        if (c.fl & TF_STRIP) {
            R ps[4];
#pragma unroll
            for (int i = 0; i < 4; i++) ps[i] = (kx[i] != R(1) || kz != R(1)) ? nx[i] : R(0);
            stg4(c.psi1 + gi, pack<R>(ps));
        }
    }
}

template <typename R, int TH>
struct AdjLoads {
    using M = EMap<TH>;
    static constexpr int EH = M::EH, AZH = TH + 16, RO = TH / 16;
    Stage<R, EH, AXW / 4> slx, sux;
    Stage<R, AZH, EW / 4> slz, suz;
    V4<R> lp[M::RMAX], up[M::RMAX], mm[M::RMAX], slp, sup, smm, gr[RO];

    __device__ __forceinline__ void load(const AdjArgs<R> &a, int x0, int z0, size_t sb, int tid)
    {
        const int ld = a.ld, q = tid & 15, g = tid >> 4;
        slx.load(a.LAx0 + sb, ld, z0 - 4, x0 - 8, tid);
        sux.load(a.UAx0 + sb, ld, z0 - 4, x0 - 8, tid);
        slz.load(a.LAz0 + sb, ld, z0 - 8, x0 - 4, tid);
        suz.load(a.UAz0 + sb, ld, z0 - 8, x0 - 4, tid);
        const R *LP0 = a.LP0 + sb, *U0 = a.U0 + sb, *cm = a.cm;
        const int er0 = M::start(g);
#pragma unroll
        for (int j = 0; j < M::RMAX; j++) {
            const int er = min(er0 + j, EH - 1);
            const unsigned gi = (unsigned)((z0 - 4 + er + ZOFF) * ld + x0 + 4 * q + XOFF);
            lp[j] = ld4(LP0 + gi);
            up[j] = ld4(U0 + gi);
            mm[j] = ld4(cm + gi);
        }
        const int sidx = tid - (NTHREADS - M::NSIDE);
        const int s_er = 4 + (sidx >= 0 ? sidx % TH : 0);
        const int s_q = (sidx >= TH) ? 17 : 0;
        const unsigned gi = (unsigned)((z0 - 4 + s_er + ZOFF) * ld + x0 - 4 + 4 * s_q + XOFF);
        slp = ld4(LP0 + gi);
        sup = ld4(U0 + gi);
        smm = ld4(cm + gi);
        const R *G = a.grad + sb;
        const int or0 = g * RO;
#pragma unroll
        for (int j = 0; j < RO; j++)
            gr[j] = ld4(G + (unsigned)((z0 + or0 + j + ZOFF) * ld + x0 + 4 * q + XOFF));
    }
    template <int W1, int W2>
    __device__ __forceinline__ void store(R (*lax)[W1], R (*uax)[W1], R (*laz)[W2], R (*uaz)[W2],
                                          int tid) const
    {
        slx.store(lax, tid);
        sux.store(uax, tid);
        slz.store(laz, tid);
        suz.store(uaz, tid);
    }
};

template <typename R, int TH>
__device__ __forceinline__ void adj_compute(const AdjArgs<R> &a, const TileId &T,
                                            const AdjLoads<R, TH> &L, int tid, R (*sLAx)[AXW],
                                            R (*sLAz)[EW], R (*sLP)[EW], R (*sUAx)[AXW],
                                            R (*sUAz)[EW], R (*sU)[EW])
{
    using M = EMap<TH>;
    constexpr int EH = M::EH, AZH = TH + 16, RO = TH / 16;
    (void)EH;
    const int ld = a.ld, q = tid & 15, g = tid >> 4;
    const int x0 = T.x0, z0 = T.z0;
    AdjCtx<R> c{a, T.s, a.X, a.Z, ld, a.T - 1 - a.t, a.tflags[T.tl], a.LPx0 + T.sb,
                a.LPx1 + T.sb, a.LP1 + T.sb, a.U1 + T.sb, a.evt};
    const int er0 = M::start(g), ern = M::count(g);
    const int sidx = tid - (NTHREADS - M::NSIDE);
    const bool has_side = sidx >= 0;
    const int s_er = 4 + (has_side ? sidx % TH : 0);
    const int s_q = (has_side && sidx >= TH) ? 17 : 0;
    const int or0 = g * RO;

    {
        const int Q = q + 1;
        V4<R> wl[8], wu[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            wl[j] = ld4(&sLAz[min(er0 + j, AZH - 1)][4 * Q]);
            wu[j] = ld4(&sUAz[min(er0 + j, AZH - 1)][4 * Q]);
        }
#pragma unroll
        for (int j = 0; j < M::RMAX; j++) {
            if (j < ern) {
                const int er = er0 + j;
                R v[12], dLx[4], dLz[4], dUx[4], dUz[4], lo[4], uo[4];
                row12<R>(&sLAx[er][4 * Q], v);
                stag_row<R, 0>(v, dLx);
                row12<R>(&sUAx[er][4 * Q], v);
                stag_row<R, 0>(v, dUx);
                stag_col<R>(wl, dLz);
                stag_col<R>(wu, dUz);
                lu_quad<R>(c, x0 + 4 * q, z0 - 4 + er, er >= 4 && er < TH + 4, dLx, dLz, dUx, dUz,
                           L.lp[j], L.up[j], L.mm[j], lo, uo);
                st4(&sLP[er][4 * Q], pack<R>(lo));
                st4(&sU[er][4 * Q], pack<R>(uo));
            }
            if (j + 1 < M::RMAX) {
#pragma unroll
                for (int k = 0; k < 7; k++) { wl[k] = wl[k + 1]; wu[k] = wu[k + 1]; }
                wl[7] = ld4(&sLAz[min(er0 + j + 8, AZH - 1)][4 * Q]);
                wu[7] = ld4(&sUAz[min(er0 + j + 8, AZH - 1)][4 * Q]);
            }
        }
        if (has_side) {
            R v[12], dLx[4], dLz[4], dUx[4], dUz[4], lo[4], uo[4];
            row12<R>(&sLAx[s_er][4 * s_q], v);
            stag_row<R, 0>(v, dLx);
            row12<R>(&sUAx[s_er][4 * s_q], v);
            stag_row<R, 0>(v, dUx);
            V4<R> ws[8];
#pragma unroll
            for (int j = 0; j < 8; j++) ws[j] = ld4(&sLAz[s_er + j][4 * s_q]);
            stag_col<R>(ws, dLz);
#pragma unroll
            for (int j = 0; j < 8; j++) ws[j] = ld4(&sUAz[s_er + j][4 * s_q]);
            stag_col<R>(ws, dUz);
            lu_quad<R>(c, x0 - 4 + 4 * s_q, z0 - 4 + s_er, false, dLx, dLz, dUx, dUz, L.slp, L.sup,
                       L.smm, lo, uo);
            st4(&sLP[s_er][4 * s_q], pack<R>(lo));
            st4(&sU[s_er][4 * s_q], pack<R>(uo));
        }
    }
    lds_barrier();

    {
        const int x = x0 + 4 * q;
        R kx[4] = {R(1), R(1), R(1), R(1)};
        if (c.fl & TF_STRIP) unpack<R>(ld4(a.kx + (unsigned)(XOFF + x)), kx);
        // lambda window: rows orr+1..orr+8; U window: rows orr..orr+8
        V4<R> wl[8], wu[9];
#pragma unroll
        for (int j = 0; j < 8; j++) wl[j] = ld4(&sLP[or0 + 1 + j][4 * q + 4]);
#pragma unroll
        for (int j = 0; j < 9; j++) wu[j] = ld4(&sU[or0 + j][4 * q + 4]);
        R *LAx1 = a.LAx1 + T.sb, *LAz1 = a.LAz1 + T.sb, *UAx1 = a.UAx1 + T.sb,
          *UAz1 = a.UAz1 + T.sb, *G = a.grad + T.sb;
        constexpr R inv5040 = R(1) / R(5040);
#pragma unroll
        for (int j = 0; j < RO; j++) {
            const int orr = or0 + j, z = z0 + orr;
            R v[12], dLx[4], dLz[4], dUx[4], dUz[4], lx[4], lz[4];
            row12<R>(&sLP[orr + 4][4 * q], v);
            stag_row<R, 1>(v, dLx);
            row12<R>(&sU[orr + 4][4 * q], v);
            stag_row<R, 1>(v, dUx);
#pragma unroll
            for (int i = 0; i < 4; i++)
                lx[i] = lap9<R>(v[i], v[i + 1], v[i + 2], v[i + 3], v[i + 4], v[i + 5], v[i + 6],
                                v[i + 7], v[i + 8]);
            stag_col<R>(wl, dLz);
            {
                V4<R> w8[8];
#pragma unroll
                for (int k = 0; k < 8; k++) w8[k] = wu[k + 1];
                stag_col<R>(w8, dUz);
                R u[9][4];
#pragma unroll
                for (int k = 0; k < 9; k++) unpack<R>(wu[k], u[k]);
#pragma unroll
                for (int i = 0; i < 4; i++)
                    lz[i] = lap9<R>(u[0][i], u[1][i], u[2][i], u[3][i], u[4][i], u[5][i], u[6][i],
                                    u[7][i], u[8][i]);
            }
            R lam[4], lax[4], laz[4], uax[4], uaz[4], gg[4];
            unpack<R>(ld4(&sLP[orr + 4][4 * q + 4]), lam);
            unpack<R>(ld4(&sLAx[orr + 4][4 * q + 8]), lax);
            unpack<R>(ld4(&sLAz[orr + 8][4 * q + 4]), laz);
            unpack<R>(ld4(&sUAx[orr + 4][4 * q + 8]), uax);
            unpack<R>(ld4(&sUAz[orr + 8][4 * q + 4]), uaz);
            unpack<R>(L.gr[j], gg);
            if (j + 1 < RO) {
#pragma unroll
                for (int k = 0; k < 7; k++) wl[k] = wl[k + 1];
                wl[7] = ld4(&sLP[orr + 9][4 * q + 4]);
#pragma unroll
                for (int k = 0; k < 8; k++) wu[k] = wu[k + 1];
                wu[8] = ld4(&sU[orr + 9][4 * q + 4]);
            }
            const R kz = (c.fl & TF_STRIP) ? a.kz[ZOFF + z] : R(1);
            R o1[4], o2[4], o3[4], o4[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                o1[i] = fma(a.cA, dLx[i], mul_rn(lax[i], kx[i]));
                o2[i] = fma(a.cA, dLz[i], mul_rn(laz[i], kz));
                o3[i] = fma(-a.cA, dUx[i], uax[i]);
                o4[i] = fma(-a.cA, dUz[i], uaz[i]);
                gg[i] = fma(lam[i], (lx[i] + lz[i]) * inv5040, gg[i]);
            }
            if (c.fl & TF_EDGE) {
#pragma unroll
                for (int i = 0; i < 4; i++)
                    if (!in_box(x + i, z, c.X, c.Z)) { o1[i] = o2[i] = o3[i] = o4[i] = R(0); }
            }
            const unsigned gi = (unsigned)((z + ZOFF) * ld + x + XOFF);
            stg4(LAx1 + gi, pack<R>(o1));
            stg4(LAz1 + gi, pack<R>(o2));
            stg4(UAx1 + gi, pack<R>(o3));
            stg4(UAz1 + gi, pack<R>(o4));
            stg4(G + gi, pack<R>(gg));
        }
    }
}

#define ADJ_LDS(TH)                                                    \
    __shared__ __align__(16) R sLAx[EMap<TH>::EH][AXW];                \
    __shared__ __align__(16) R sLAz[TH + 16][EW];                      \
    __shared__ __align__(16) R sLP[EMap<TH>::EH][EW];                  \
    __shared__ __align__(16) R sUAx[EMap<TH>::EH][AXW];                \
    __shared__ __align__(16) R sUAz[TH + 16][EW];                      \
    __shared__ __align__(16) R sU[EMap<TH>::EH][EW]

template <typename R, int TH>
__global__ void __launch_bounds__(NTHREADS)
adj_step_kernel(const AdjArgs<R> a)
{
    static_assert(TH % 16 == 0, "tile height must be a multiple of 16");
    ADJ_LDS(TH);
    const int tid = threadIdx.x;
    const TileId T = tile_id<TH>(xcd_remap(blockIdx.x, gridDim.x), a.ntx, a.ntz, a.ld, a.sstride);
    AdjLoads<R, TH> L;
    L.load(a, T.x0, T.z0, T.sb, tid);
    L.store(sLAx, sUAx, sLAz, sUAz, tid);
    lds_barrier();
    adj_compute<R, TH>(a, T, L, tid, sLAx, sLAz, sLP, sUAx, sUAz, sU);
}

template <typename R, int TH>
__global__ void __launch_bounds__(NTHREADS)
adj_step_persistent(const AdjArgs<R> a)
{
    ADJ_LDS(TH);
    const int tid = threadIdx.x;
    const int nb = a.ntx * a.ntz * a.nshots;
    const int xcd = blockIdx.x & 7, k = blockIdx.x >> 3;
    const int nk = (gridDim.x - xcd + 7) >> 3;
    int lo, hi;
    xcd_chunk(xcd, nb, lo, hi);
    int Lt = lo + k;
    AdjLoads<R, TH> L;
    TileId T = tile_id<TH>(min(Lt, nb - 1), a.ntx, a.ntz, a.ld, a.sstride);
    if (Lt < hi) L.load(a, T.x0, T.z0, T.sb, tid);
    for (; Lt < hi; Lt += nk) {
        L.store(sLAx, sUAx, sLAz, sUAz, tid);
        AdjLoads<R, TH> C;
#pragma unroll
        for (int j = 0; j < EMap<TH>::RMAX; j++) { C.lp[j] = L.lp[j]; C.up[j] = L.up[j]; C.mm[j] = L.mm[j]; }
#pragma unroll
        for (int j = 0; j < TH / 16; j++) C.gr[j] = L.gr[j];
        C.slp = L.slp;
        C.sup = L.sup;
        C.smm = L.smm;
        const TileId Tc = T;
        lds_barrier();
        if (Lt + nk < hi) {
            T = tile_id<TH>(Lt + nk, a.ntx, a.ntz, a.ld, a.sstride);
            L.load(a, T.x0, T.z0, T.sb, tid);
        }
        adj_compute<R, TH>(a, Tc, C, tid, sLAx, sLAz, sLP, sUAx, sUAz, sU);
        lds_barrier();
    }
}

template <typename R, int TH>
static hipError_t fwd_launch_th(const FwdArgs<R> &a, int persistent_wgs, hipStream_t st)
{
    const int nb = a.ntx * a.ntz * a.nshots;
    if (persistent_wgs > 0) {
        hipLaunchKernelGGL((fwd_step_persistent<R, TH>), dim3(min(persistent_wgs, nb)),
                           dim3(NTHREADS), 0, st, a);
    } else {
        hipLaunchKernelGGL((fwd_step_kernel<R, TH>), dim3(nb), dim3(NTHREADS), 0, st, a);
    }
    return hipGetLastError();
}

template <typename R, int TH>
static hipError_t adj_launch_th(const AdjArgs<R> &a, int persistent_wgs, hipStream_t st)
{
    const int nb = a.ntx * a.ntz * a.nshots;
    if (persistent_wgs > 0) {
        hipLaunchKernelGGL((adj_step_persistent<R, TH>), dim3(min(persistent_wgs, nb)),
                           dim3(NTHREADS), 0, st, a);
    } else {
        hipLaunchKernelGGL((adj_step_kernel<R, TH>), dim3(nb), dim3(NTHREADS), 0, st, a);
    }
    return hipGetLastError();
}

template <typename R, int TH>
static int occ_fwd(bool persistent)
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &n, persistent ? (const void *)fwd_step_persistent<R, TH> : (const void *)fwd_step_kernel<R, TH>,
            NTHREADS, 0) != hipSuccess)
        return 0;
    return n;
}
template <typename R, int TH>
static int occ_adj(bool persistent)
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &n, persistent ? (const void *)adj_step_persistent<R, TH> : (const void *)adj_step_kernel<R, TH>,
            NTHREADS, 0) != hipSuccess)
        return 0;
    return n;
}

template <>
hipError_t launch_fwd_step<float>(const FwdArgs<float> &a, int th, int pw, hipStream_t s)
{
    switch (th) {
    case 16: return fwd_launch_th<float, 16>(a, pw, s);
    case 32: return fwd_launch_th<float, 32>(a, pw, s);
    case 64: return fwd_launch_th<float, 64>(a, pw, s);
    default: return hipErrorInvalidValue;
    }
}
template <>
hipError_t launch_fwd_step<double>(const FwdArgs<double> &a, int th, int pw, hipStream_t s)
{
    switch (th) {
    case 16: return fwd_launch_th<double, 16>(a, pw, s);
    case 32: return fwd_launch_th<double, 32>(a, pw, s);
    default: return hipErrorInvalidValue;
    }
}
template <>
hipError_t launch_adj_step<float>(const AdjArgs<float> &a, int th, int pw, hipStream_t s)
{
    switch (th) {
    case 16: return adj_launch_th<float, 16>(a, pw, s);
    case 32: return adj_launch_th<float, 32>(a, pw, s);
    case 64: return adj_launch_th<float, 64>(a, pw, s);
    default: return hipErrorInvalidValue;
    }
}
template <>
hipError_t launch_adj_step<double>(const AdjArgs<double> &a, int th, int pw, hipStream_t s)
{
    switch (th) {
    case 16: return adj_launch_th<double, 16>(a, pw, s);
    default: return hipErrorInvalidValue;
    }
}

// ---------------------------------------------------------------------------
// two time steps per launch: z-marching strips
// ---------------------------------------------------------------------------
// A work-group of NW = 256 or 384 threads owns NW - 32 columns
// [xs, xs + NW - 32) of one shot (thread tid holds column xs - F2_HALO + tid)
// and output rows
// [zc0, zc0 + F2_CH), and walks down the rows.  Iteration j takes row
// i = zc0 - 15 + j of the state before the pair and computes, each stage on
// rows already complete below it,
//   P(t) on row i - 3,  A(t) on row i - 7,  P(t+1) on row i - 10 (written
//   out),  A(t+1) on row i - 14 (written out).
// z-stencils read per-thread register rings of the last 8 rows (a stage
// that needs 9 rows reads the slot about to be overwritten); x-stencils and
// the delayed A values read LDS rings of 8 rows written in earlier
// iterations, so one LDS barrier per iteration orders everything.  The
// 16-column halo shrinks by at most 4 columns per stage and ends 1 column
// outside the owned ones.  Global loads run f2_pf rows ahead in registers.
// Every cell is updated with the arithmetic of the one-step kernels, so
// results are bitwise equal; the state before the pair is read from one slot
// and the state after it written to the other.
//
// Three field kinds share the scheme (MODE):
//   F2_FWD  forward field: damped operator, source injection (icum),
//           receiver and boundary-ring sampling (fwd_step_kernel);
//   F2_LAM  adjoint field lambda: damped operator, injection of the running
//           sum of the adjoint source at the receivers (jcum), and
//           grad += lambda_P * lap of the reconstructed field (read from
//           the lap planes written by F2_REC) (adj_step_kernel's lambda);
//   F2_REC  reconstructed forward field U: undamped reverse operator,
//           boundary-ring values written back, and lap(U) (x plus z
//           9-point second differences / 5040) written for both steps
//           (adj_step_kernel's U).
// Memory is accessed through buffer resources: an offset at or beyond the
// buffer's size reads 0 and drops a store without touching memory, so rows
// and columns outside the plane, undamped psi cells, cells without an event
// and lanes that do not own a cell need no branch.  An offset is a lane part
// (column) plus a row part, either of which can rule the access out (F2_SENT).
// In the forward kernel the iterations whose rows all lie inside the chunk
// and away from the plane's edges run a copy of the body without row checks.
constexpr int F2_FWD = 0, F2_LAM = 1, F2_REC = 2;
// constant address space: uniform loads through it are scalar and invariant
#define F2_CONST __attribute__((address_space(4)))
// Rows of global loads in flight ahead of the row being computed, per mode
// (measured at 64 x 1024^2: the reconstruction kernel gains 8 % at 4, the
// forward kernel loses 90 % there and the lambda kernel nothing).
#ifndef F2_PREFETCH_REC
#define F2_PREFETCH_REC 4
#endif
#ifndef F2_PREFETCH_LAM
#define F2_PREFETCH_LAM 2
#endif
#ifndef F2_PREFETCH
#define F2_PREFETCH 2
#endif
#ifndef F2_BARRIER
#define F2_BARRIER 2
#endif
static_assert(F2_BARRIER == 1 || F2_BARRIER == 2, "LDS ring hazards allow a barrier every 1 or 2 rows");
#ifndef F2_CHUNK
#define F2_CHUNK 256
#endif
constexpr int F2_CH = F2_CHUNK;   // output rows per work-group
template <int MODE>
constexpr int f2_pf() { return MODE == 2 ? F2_PREFETCH_REC : MODE == 1 ? F2_PREFETCH_LAM : F2_PREFETCH; }
// Per-mode code shape, bit MODE of each mask (measured at 64 x 1024^2, ms per
// launch of the forward / reconstruction / lambda kernel):
//  F2_STEADY_MODES: iterations in the middle of a chunk run a second copy of
//    the 8-row body without row checks (254 instead of 357 instructions per
//    row in the forward kernel: 0.645 -> 0.570 ms; the reconstruction and
//    lambda kernels ran 6-8 % slower with it, at more registers).
//  F2_OPAQUE_MODES: the row part of an offset is pinned in a scalar register
//    (see f2_row): the reconstruction kernel needs 127 vector registers with
//    it and 168 without (0.604 vs 0.644 ms); the forward kernel ran 3 %
//    slower with it.
#ifndef F2_STEADY_MODES
#define F2_STEADY_MODES 1
#endif
#ifndef F2_OPAQUE_MODES
#define F2_OPAQUE_MODES 6
#endif
template <int MODE>
constexpr bool f2_steady() { return ((F2_STEADY_MODES) >> MODE) & 1; }
template <int MODE>
constexpr bool f2_opaque() { return ((F2_OPAQUE_MODES) >> MODE) & 1; }
// Block geometry for NW threads: owned columns, LDS row width (8 padding
// cells each side), ring-table words per row (slot base, one 16-bit prefix
// per 32-lane word, one lane mask per word).
template <int NW>
struct F2Geo {
    static_assert(NW == 128 || NW == 192 || NW == 256 || NW == 384,
                  "two-step blocks are 128, 192, 256 or 384 threads");
    static constexpr int OWN = NW - 2 * F2_HALO, LP = NW + 16;
    static constexpr int WORDS = NW / 32, RW = 1 + NW / 64 + NW / 32;
};
// LDS row tables (kz, lane masks) cover rows [zc0 - F2_RB, zc0 - F2_RB + F2_NR):
// iterations touch rows zc0 - 29 .. zc0 + F2_CH + 20 + prefetch (<= 4)
constexpr int F2_RB = 32, F2_NR = F2_CH + 64;
constexpr int F2_LEAD = 15;   // rows loaded above a chunk
// Buffer offsets are formed as a lane part (the column; loop-invariant, in a
// vector register) plus a row part (row, slot and time offsets; the same for
// every lane, in a scalar register), each F2_SENT where that part rules the
// access out.  Every resource spans at most F2_SENT bytes, so a sum holding
// F2_SENT is out of range, and 2 * F2_SENT does not wrap: no per-access
// select is needed.
constexpr unsigned F2_SENT = 0x7FFFFFF0u;

typedef __amdgpu_buffer_rsrc_t Rsrc;

__device__ __forceinline__ Rsrc f2_rsrc(const void *base, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes,
                                             0x00020000);
}
template <typename R> __device__ __forceinline__ R bload(Rsrc r, unsigned off);
template <> __device__ __forceinline__ float bload<float>(Rsrc r, unsigned off)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
template <> __device__ __forceinline__ double bload<double>(Rsrc r, unsigned off)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
template <typename R> __device__ __forceinline__ void bstore(Rsrc r, unsigned off, R v);
template <> __device__ __forceinline__ void bstore<float>(Rsrc r, unsigned off, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0, 0);
}
template <> __device__ __forceinline__ void bstore<double>(Rsrc r, unsigned off, double v)
{
    typedef unsigned U2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, v), r, off, 0, 0);
}

// LDS row segment ring[q][col .. col + N).  (Hiding the row offset from the
// compiler, so that it forms one address per row instead of one per pair of
// reads, cut 15 vector instructions per iteration but made all three kernels
// 2-5 % slower on MI355X; the plain form is kept.)
template <typename R, int N, int LP>
__device__ __forceinline__ void lds_seg(R (*ring)[LP], int q, int col, R (&v)[N])
{
#pragma unroll
    for (int k = 0; k < N; k++) v[k] = ring[q][col + k];
}

template <typename R, int LP>
__device__ __forceinline__ R stag_lds(R (*ring)[LP], int q, int c0)
{
    R v[8];
    lds_seg<R, 8, LP>(ring, q, c0, v);
    return stag<R>(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]);
}

// Event of one cell, each index + 1 (0: none), as the per-cell event map holds it.
struct F2Evt {
    int ring1, rx1, src1;
};

// What iteration i needs from global memory, loaded f2_pf iterations ahead.
template <typename R, int MODE>
struct F2Row {
    R ax, az, p, ps, m;   // state before the pair and m on row i (ps: FWD, LAM)
    R m3;                 // m on row i - 10
    F2Evt e1, e3;         // FWD, LAM: events on rows i - 3 and i - 10
    R g, l0, l1;          // LAM: grad and lap(U(t)) on row i - 7, lap(U(t+1)) on row i - 14
    unsigned k1, k3;      // FWD, REC: lane parts of the ring slots on rows i - 3 and
                          // i - 10 (slot * sizeof(R); F2_SENT: none)
    R r1, r3;             // REC: ring values there at steps t and t+1
};

template <typename R, int MODE>
struct Pair2Regs {
    R az0[8];                 // Az before the pair by row, newest = row i
    R p0[4], s0[4], m0[4];    // P, psi, m before the pair: rows i-3 .. i
    R p1[8], s1[8];           // P(t), psi(t): rows i-11 .. i-4 (before this iteration's write)
    R az1[8];                 // Az(t): rows i-14 .. i-7
    R p2[8];                  // P(t+1): rows i-18 .. i-11 (before this iteration's write)
    R gq[8];                  // LAM: grad after step t, rows i-14 .. i-7
    F2Row<R, MODE> pf[f2_pf<MODE>()];
};

struct F2Ctx {
    int tid, x, s, ss, zc0, zc1, lc, X, Z, ld, t, T, nring;   // ss: active source (-1: all)
    // lane parts of offsets: the byte offset of this lane's column in a
    // padded row (element size R) where the lane may access it, else F2_SENT
    unsigned loc;    // loads: columns of the block that reach an owned cell
    unsigned loo;    // stores: owned columns
    unsigned lod;    // psi loads on rows without z damping: loc where x is damped
    unsigned lodo;   // psi stores on such rows: loo where x is damped
    unsigned lo16;   // event map (16-byte cells): loads
    bool own, xin;   // xin: 4 <= x < X - 4
    // one resource per field covers both slots of the shot; the slot read and
    // the slot written are byte offsets (P kind: P, psi; A kind: Ax, Az)
    Rsrc rP, rAx, rAz, rPx, rcm, rpm, rg, rl, rring;
    unsigned dPi, dPo, dAi, dAo, dL1;   // dL1: lap1 relative to lap0 (one resource)
};

// Row part of a plane offset: row z of a padded plane plus d bytes, or
// F2_SENT when ok is false.
template <typename R, int MODE>
__device__ __forceinline__ unsigned f2_row(const F2Ctx &c, int z, unsigned d, bool ok)
{
    unsigned r = ok ? (unsigned)((z + ZOFF) * c.ld) * (unsigned)sizeof(R) + d : F2_SENT;
    // f2_opaque: the row part stays a scalar formed where it is used, which
    // keeps the compiler from re-associating lane + row and holding lane +
    // (row constant) of every unrolled row in vector registers.  Emits no
    // instruction.
    if constexpr (f2_opaque<MODE>()) asm("" : "+s"(r));
    return r;
}

__device__ __forceinline__ bool f2_zin(const F2Ctx &c, int z) { return z >= BOX && z < c.Z - BOX; }
__device__ __forceinline__ bool f2_chunk(const F2Ctx &c, int z) { return z >= c.zc0 && z < c.zc1; }

// source / receiver event of this lane's cell on row z (FWD, LAM): sev
// flags the rows holding one in the block's columns; on those rows every
// lane reads its cell of the event map (a few rows per block).  The flag is
// the same for every lane, so it is read into a scalar.
__device__ __forceinline__ F2Evt f2_evt(const F2Ctx &c, const uint8_t *sev, int z)
{
    const int row = __builtin_amdgcn_readfirstlane((int)sev[z - c.zc0 + F2_RB]);
    const unsigned off = c.lo16 + (row != 0 ? (unsigned)((z + ZOFF) * c.ld) * 16u : F2_SENT);
    const auto v = __builtin_amdgcn_raw_buffer_load_b96(c.rpm, off, 0, 0);
    return F2Evt{(int)v[0], (int)v[1], (int)v[2]};
}

// Lane part of the ring slot of this lane's cell on row z (FWD, REC): slot *
// es, or F2_SENT when the cell is not a ring cell or allow is false.  Per
// row smask holds F2Geo<NW>::RW words: the ring slot of the block's first
// ring cell, the ring cells before each 32-lane word (16 bits each), and the
// lane masks; ring slots are numbered in raster order.
template <int NW>
__device__ __forceinline__ unsigned f2_rslot(const F2Ctx &c, const unsigned *smask, int z,
                                             unsigned es, bool allow)
{
    const int w = c.tid >> 5, b = c.tid & 31;
    const unsigned *e = smask + (z - c.zc0 + F2_RB) * F2Geo<NW>::RW;
    const unsigned mk = e[1 + NW / 64 + w];
    const unsigned pre = (e[1 + (w >> 1)] >> (16 * (w & 1))) & 0xFFFFu;
    const unsigned k = e[0] + pre + (unsigned)__builtin_popcount(mk & ((1u << b) - 1u));
    return allow && ((mk >> b) & 1u) ? k * es : F2_SENT;
}

// ST (steady): the caller guarantees that every row this iteration touches
// lies in the plane, in the work-group's chunk and 4 or more rows from the
// plane's top and bottom edges, so no row check is needed.
template <typename R, int MODE, int NW, bool ST>
__device__ __forceinline__ void f2_fetch(const F2Ctx &c, const unsigned *smask, const uint8_t *sev,
                                         const R *skz, int i, F2Row<R, MODE> &f)
{
    const bool rv = ST || (i >= -ZOFF && i < c.Z + 4);
    const unsigned sA = f2_row<R, MODE>(c, i, c.dAi, rv), sP = f2_row<R, MODE>(c, i, c.dPi, rv);
    f.ax = bload<R>(c.rAx, c.loc + sA);
    f.az = bload<R>(c.rAz, c.loc + sA);
    f.p = bload<R>(c.rP, c.loc + sP);
    f.m = bload<R>(c.rcm, c.loc + f2_row<R, MODE>(c, i, 0u, rv));
    const int i3 = i - 10;
    f.m3 = bload<R>(c.rcm, c.loc + f2_row<R, MODE>(c, i3, 0u, ST || (i3 >= -ZOFF && i3 < c.Z + 4)));
    f.ps = f.g = f.l0 = f.l1 = f.r1 = f.r3 = R(0);
    f.k1 = f.k3 = F2_SENT;
    f.e1 = f.e3 = F2Evt{0, 0, 0};
    if constexpr (MODE != F2_REC) {
        // psi is stored on damped cells only: x-damped columns, or every
        // column of a z-damped row
        const bool zd = skz[i - c.zc0 + F2_RB] != R(1);
        f.ps = bload<R>(c.rPx, (zd ? c.loc : c.lod) + sP);
        f.e1 = f2_evt(c, sev, i - 3);
        f.e3 = f2_evt(c, sev, i3);
    }
    if constexpr (MODE != F2_LAM) {
        const unsigned es = (unsigned)sizeof(R);
        const bool allow = MODE == F2_REC || c.own;   // the forward samples owned cells
        f.k1 = f2_rslot<NW>(c, smask, i - 3, es, allow);
        f.k3 = f2_rslot<NW>(c, smask, i3, es, allow);
    }
    if constexpr (MODE == F2_LAM) {
        const int z2 = i - 7, z4 = i - 14;
        const unsigned s2 = f2_row<R, MODE>(c, z2, 0u, ST || f2_chunk(c, z2));
        f.g = bload<R>(c.rg, c.loo + s2);
        f.l0 = bload<R>(c.rl, c.loo + s2);
        f.l1 = bload<R>(c.rl, c.loo + f2_row<R, MODE>(c, z4, c.dL1, ST || f2_chunk(c, z4)));
    }
    if constexpr (MODE == F2_REC) {
        // ring samples of the shot, time-major: ring times T-1-t and T-2-t
        const unsigned es = (unsigned)sizeof(R);
        const int r1 = c.T - 1 - c.t, r3 = c.T - 2 - c.t;
        f.r1 = bload<R>(c.rring, f.k1 + (r1 >= 0 ? (unsigned)(r1 * c.nring) * es : F2_SENT));
        f.r3 = bload<R>(c.rring, f.k3 + (r3 >= 0 ? (unsigned)(r3 * c.nring) * es : F2_SENT));
    }
}

// P-kind update of one cell at step tt (p_quad / lu_quad arithmetic for one
// lane).  rec_here = this block samples the cell (FWD).
// ks: lane part of the cell's ring slot (F2_SENT: none): REC writes the ring
// value rv back, FWD samples P there at ks + sring (sring: the row part,
// F2_SENT outside the chunk), both branch-free.
template <typename R, int MODE>
__device__ __forceinline__ R f2_p(const Pair2Args<R> &a, const F2Ctx &c, const F2Evt &e, R rv,
                                  unsigned ks, unsigned sring, bool inb, int tt, bool rec_here, R pp,
                                  R psi, R mm, R kx, R kz, R dAx, R dAz, R &psi_out)
{
    if constexpr (MODE == F2_REC) {
        R up = fma(-mm, dAx + dAz, pp);
        if (ks != F2_SENT) up = rv;
        psi_out = R(0);
        return inb ? up : R(0);
    } else {
        R nx = fma(mm, dAx, mul_rn(psi, kx));
        R nz = fma(mm, dAz, mul_rn(pp - psi, kz));
        if ((e.rx1 | e.src1) != 0) {
            if constexpr (MODE == F2_FWD) {
                const int src = e.src1 - 1;
                if (src >= 0 && (c.ss < 0 || c.ss == src)) {
                    const R h = a.icum[(size_t)src * a.T + tt] * R(0.5);
                    nx = add_rn(nx, h);
                    nz = add_rn(nz, h);
                }
                if (rec_here && e.rx1 > 0 && a.rec)
                    a.rec[((size_t)c.s * a.n_rx + (e.rx1 - 1)) * a.T + tt] = add_rn(nx, nz);
            } else {
                if (e.rx1 > 0) {
                    const R h = a.jcum[((size_t)c.s * a.n_rx + (e.rx1 - 1)) * a.T + tt] * R(0.5);
                    nx = add_rn(nx, h);
                    nz = add_rn(nz, h);
                }
            }
        }
        R p = add_rn(nx, nz);
        if constexpr (MODE == F2_FWD)   // ring cells lie inside the box: p is the sample
            bstore<R>(c.rring, ks + sring, p);
        if (!inb) { p = R(0); nx = R(0); }
        psi_out = (kx != R(1) || kz != R(1)) ? nx : R(0);
        return p;
    }
}

// A-kind update of one cell
template <typename R, int MODE>
__device__ __forceinline__ void f2_a(R cA, R dPx, R dPz, R ax, R az, R kx, R kz, bool inb, R &nax,
                                     R &naz)
{
    if constexpr (MODE == F2_REC) {
        nax = fma(-cA, dPx, ax);
        naz = fma(-cA, dPz, az);
    } else {
        nax = fma(cA, dPx, mul_rn(ax, kx));
        naz = fma(cA, dPz, mul_rn(az, kz));
    }
    if (!inb) { nax = R(0); naz = R(0); }
}

// (lap9 along x + lap9 along z) / 5040 as adj_step_kernel forms it: row =
// the LDS row (9 taps around lc), w = the 9 rows (z-4 .. z+4) of this column
template <typename R, int LP>
__device__ __forceinline__ R f2_lap(R (*ring)[LP], int q, int lc, const R (&w)[9])
{
    constexpr R inv5040 = R(1) / R(5040);
    R row[9];
    lds_seg<R, 9, LP>(ring, q, lc - 4, row);
    const R lx = lap9<R>(row[0], row[1], row[2], row[3], row[4], row[5], row[6], row[7], row[8]);
    const R lz = lap9<R>(w[0], w[1], w[2], w[3], w[4], w[5], w[6], w[7], w[8]);
    return (lx + lz) * inv5040;
}

template <typename R, int MODE, int NW, int PH, bool ST>
__device__ __forceinline__ void f2_iter(const Pair2Args<R> &a, const F2Ctx &c,
                                        Pair2Regs<R, MODE> &g, R (*sAx0)[F2Geo<NW>::LP],
                                        R (*sP1)[F2Geo<NW>::LP], R (*sAx1)[F2Geo<NW>::LP],
                                        R (*sP2)[F2Geo<NW>::LP],
                                        const unsigned *smask, const uint8_t *sev, const R *skz,
                                        int i, R kx)
{
    const unsigned es = (unsigned)sizeof(R);
    constexpr int Q1 = (PH + 1) & 7, Q4 = (PH + 4) & 7, Q5 = (PH + 5) & 7;
    constexpr int F = PH % f2_pf<MODE>();
    constexpr int H = PH & 3, H1 = (PH + 1) & 3;   // 4-row rings: newest, 3 rows back
    const int lc = c.lc;
    auto kz_at = [&](int z) -> R {
        if constexpr (MODE == F2_REC) return R(1);
        else return skz[z - c.zc0 + F2_RB];
    };
    F2Row<R, MODE> &f = g.pf[F];
    // ---- 1. take row i of the state before the pair ----
    g.az0[PH] = f.az; g.p0[H] = f.p; g.m0[H] = f.m;
    if constexpr (MODE != F2_REC) g.s0[H] = f.ps;
    sAx0[PH][lc] = f.ax;
    // ---- 2. P(t) on row z1 = i - 3 ----
    const int z1 = i - 3;
    R pn1, sn1;
    {
        const R dAx = stag_lds<R>(sAx0, Q5, lc - 4);
        const R dAz = stag<R>(g.az0[(PH + 1) & 7], g.az0[(PH + 2) & 7], g.az0[(PH + 3) & 7],
                              g.az0[(PH + 4) & 7], g.az0[(PH + 5) & 7], g.az0[(PH + 6) & 7],
                              g.az0[(PH + 7) & 7], g.az0[PH]);
        const bool in1 = ST || f2_chunk(c, z1);
        const unsigned sring = in1 ? (unsigned)(c.t * c.nring) * es : F2_SENT;
        pn1 = f2_p<R, MODE>(a, c, f.e1, f.r1, f.k1, sring, c.xin && (ST || f2_zin(c, z1)), c.t,
                            c.own && in1, g.p0[H1], g.s0[H1], g.m0[H1], kx, kz_at(z1), dAx, dAz,
                            sn1);
        sP1[PH][lc] = pn1;
    }
    // ---- 3. A(t) on row z2 = i - 7 ----
    {
        const int z2 = i - 7;
        // P(t) rows z2-4 .. z2+4: p1[PH] .. p1[PH+7] hold rows i-11 .. i-4
        const R dPx = stag_lds<R>(sP1, Q4, lc - 3);
        const R dPz = stag<R>(g.p1[(PH + 1) & 7], g.p1[(PH + 2) & 7], g.p1[(PH + 3) & 7],
                              g.p1[(PH + 4) & 7], g.p1[(PH + 5) & 7], g.p1[(PH + 6) & 7],
                              g.p1[(PH + 7) & 7], pn1);
        R nax, naz;
        f2_a<R, MODE>(a.cA, dPx, dPz, sAx0[Q1][lc], g.az0[Q1], kx, kz_at(z2),
                      c.xin && (ST || f2_zin(c, z2)), nax, naz);
        g.az1[PH] = naz;
        sAx1[PH][lc] = nax;
        if constexpr (MODE == F2_LAM) g.gq[PH] = fma(g.p1[Q4], f.l0, f.g);
        if constexpr (MODE == F2_REC) {
            const R w[9] = {g.p1[PH], g.p1[(PH + 1) & 7], g.p1[(PH + 2) & 7], g.p1[(PH + 3) & 7],
                            g.p1[(PH + 4) & 7], g.p1[(PH + 5) & 7], g.p1[(PH + 6) & 7],
                            g.p1[(PH + 7) & 7], pn1};
            const R lp = f2_lap<R>(sP1, Q4, lc, w);
            bstore<R>(c.rl, (ST || f2_chunk(c, z2)) ? c.loo + f2_row<R, MODE>(c, z2, 0u, true) : F2_SENT, lp);
        }
        g.p1[PH] = pn1;
        if constexpr (MODE != F2_REC) g.s1[PH] = sn1;
    }
    // ---- 4. P(t+1) on row z3 = i - 10, written out ----
    const int z3 = i - 10;
    R pn2;
    {
        const R dAx = stag_lds<R>(sAx1, Q5, lc - 4);
        const R dAz = stag<R>(g.az1[(PH + 1) & 7], g.az1[(PH + 2) & 7], g.az1[(PH + 3) & 7],
                              g.az1[(PH + 4) & 7], g.az1[(PH + 5) & 7], g.az1[(PH + 6) & 7],
                              g.az1[(PH + 7) & 7], g.az1[PH]);
        const R kz3 = kz_at(z3);
        R sn2;
        const bool in3 = ST || f2_chunk(c, z3);
        const unsigned sring = in3 ? (unsigned)((c.t + 1) * c.nring) * es : F2_SENT;
        pn2 = f2_p<R, MODE>(a, c, f.e3, f.r3, f.k3, sring, c.xin && (ST || f2_zin(c, z3)),
                            c.t + 1, c.own && in3, g.p1[Q1], g.s1[Q1], f.m3, kx, kz3, dAx, dAz,
                            sn2);
        sP2[PH][lc] = pn2;
        const unsigned so = f2_row<R, MODE>(c, z3, c.dPo, true);
        bstore<R>(c.rP, in3 ? c.loo + so : F2_SENT, pn2);
        if constexpr (MODE != F2_REC)   // psi on damped cells
            bstore<R>(c.rPx, in3 ? (kz3 != R(1) ? c.loo : c.lodo) + so : F2_SENT, sn2);
    }
    // ---- 5. A(t+1) on row z4 = i - 14, written out ----
    {
        const int z4 = i - 14;
        // P(t+1) rows z4-4 .. z4+4: p2[PH] .. p2[PH+7] hold rows i-18 .. i-11
        const R dPx = stag_lds<R>(sP2, Q4, lc - 3);
        const R dPz = stag<R>(g.p2[(PH + 1) & 7], g.p2[(PH + 2) & 7], g.p2[(PH + 3) & 7],
                              g.p2[(PH + 4) & 7], g.p2[(PH + 5) & 7], g.p2[(PH + 6) & 7],
                              g.p2[(PH + 7) & 7], pn2);
        R nax, naz;
        f2_a<R, MODE>(a.cA, dPx, dPz, sAx1[Q1][lc], g.az1[Q1], kx, kz_at(z4),
                      c.xin && (ST || f2_zin(c, z4)), nax, naz);
        const bool in4 = ST || f2_chunk(c, z4);
        const unsigned sA = f2_row<R, MODE>(c, z4, c.dAo, true);
        bstore<R>(c.rAx, in4 ? c.loo + sA : F2_SENT, nax);
        bstore<R>(c.rAz, in4 ? c.loo + sA : F2_SENT, naz);
        if constexpr (MODE == F2_LAM)
            bstore<R>(c.rg, in4 ? c.loo + f2_row<R, MODE>(c, z4, 0u, true) : F2_SENT,
                      fma(g.p2[Q4], f.l1, g.gq[Q1]));
        if constexpr (MODE == F2_REC) {
            const R w[9] = {g.p2[PH], g.p2[(PH + 1) & 7], g.p2[(PH + 2) & 7], g.p2[(PH + 3) & 7],
                            g.p2[(PH + 4) & 7], g.p2[(PH + 5) & 7], g.p2[(PH + 6) & 7],
                            g.p2[(PH + 7) & 7], pn2};
            bstore<R>(c.rl, in4 ? c.loo + f2_row<R, MODE>(c, z4, c.dL1, true) : F2_SENT, f2_lap<R>(sP2, Q4, lc, w));
        }
        g.p2[PH] = pn2;
    }
    // ---- 6. loads for the row f2_pf ahead into the slot just consumed ----
    f2_fetch<R, MODE, NW, ST>(c, smask, sev, skz, i + f2_pf<MODE>(), f);
    // Cross-lane LDS reads are of rows written 3 or 4 iterations earlier, and
    // a row slot is rewritten 8 iterations after it was written (4 after its
    // last cross-lane read); reads of a lane's own column stay in its wave.
    // A barrier every F2_BARRIER iterations (1 or 2) orders all of them.
    if constexpr ((PH + 1) % F2_BARRIER == 0) lds_barrier();
}

// evrow (FWD, LAM): [ceil(X / (NW - 32))][Z] flags of the rows holding a
// source or receiver in each block's NW columns, with pmap the per-cell event
// map of the padded plane, int4 {ring + 1, rx + 1, src + 1, 0} (0: none);
// mask (FWD, REC): [..][Z][F2Geo<NW>::RW] ring-cell tables as f2_rslot reads.
// Waves per SIMD the register allocation targets: work-groups of 4 waves,
// three per CU in FP32 (FWD 50 KB and REC 49 KB of LDS allow 3; LAM at 4 was
// slower); FP64 (validation) leaves the allocation unconstrained.
template <typename R, int MODE, int NW>
constexpr int f2_waves() { return sizeof(R) == 4 ? 3 : 1; }

// Work-group coordinates of a two-step launch (column block x, row chunk y,
// shot z): the launch grid's own, or given.
struct F2GridIds {
    __device__ unsigned x() const { return blockIdx.x; }
    __device__ unsigned y() const { return blockIdx.y; }
    __device__ unsigned z() const { return blockIdx.z; }
};
struct F2GivenIds {
    unsigned bx, by, bz;
    __device__ unsigned x() const { return bx; }
    __device__ unsigned y() const { return by; }
    __device__ unsigned z() const { return bz; }
};

// One work-group of a two-step launch of the problem described by a.
template <typename R, int MODE, int NW, typename Ids>
__device__ __forceinline__ void pair2_body(const Pair2Args<R> a, const Ids id)
{
    using G = F2Geo<NW>;
    constexpr int LP = G::LP;
    constexpr int MW = MODE == F2_LAM ? 1 : G::RW;    // ring-table words per row
    constexpr int EV = MODE == F2_REC ? 4 : F2_NR;    // event-row flags
    __shared__ R sAx0[8][LP], sP1[8][LP], sAx1[8][LP], sP2[8][LP];
    __shared__ R skz[MODE == F2_REC ? 1 : F2_NR];
    __shared__ unsigned smask[F2_NR * MW];
    __shared__ uint8_t sev[EV];
    F2Ctx c;
    c.tid = threadIdx.x;
    c.X = a.X; c.Z = a.Z; c.ld = a.ld; c.t = a.t; c.T = a.T; c.nring = a.n_ring;
    const int xs = id.x() * G::OWN;
    c.x = xs - F2_HALO + c.tid;
    c.zc0 = id.y() * F2_CH;
    c.zc1 = min(c.zc0 + F2_CH, a.Z);
    c.s = id.z();
    c.ss = MODE == F2_FWD ? a.shot_src[c.s] : -1;
    c.lc = c.tid + 8;
    // columns beyond X + 12 never reach an owned cell: not loaded
    const bool colv = c.x < a.X + 12;
    c.own = c.tid >= F2_HALO && c.tid < F2_HALO + G::OWN && c.x < a.X;
    c.xin = c.x >= BOX && c.x < a.X - BOX;
    const unsigned col = (unsigned)(min(c.x, a.X + 11) + XOFF);
    const R kx = (MODE != F2_REC && colv) ? a.kx[XOFF + c.x] : R(1);
    c.loc = colv ? col * (unsigned)sizeof(R) : F2_SENT;
    c.loo = c.own ? col * (unsigned)sizeof(R) : F2_SENT;
    c.lod = kx != R(1) ? c.loc : F2_SENT;
    c.lodo = kx != R(1) ? c.loo : F2_SENT;
    c.lo16 = colv ? col * 16u : F2_SENT;
    const size_t sb = (size_t)c.s * (size_t)a.sstride;
    const unsigned pb = (unsigned)((size_t)a.sstride * sizeof(R));
    // both slots of a field: slot 0 + slot_elems (the lower pointer is slot 0)
    const unsigned fb = (unsigned)((size_t)(a.slot_elems + a.sstride) * sizeof(R));
    const R *bP = a.P0 < a.P1 ? a.P0 : a.P1, *bA = a.Ax0 < a.Ax1 ? a.Ax0 : a.Ax1;
    const R *bAz = a.Az0 < a.Az1 ? a.Az0 : a.Az1;
    c.dPi = (unsigned)((size_t)(a.P0 - bP) * sizeof(R));
    c.dPo = (unsigned)((size_t)(a.P1 - bP) * sizeof(R));
    c.dAi = (unsigned)((size_t)(a.Ax0 - bA) * sizeof(R));
    c.dAo = (unsigned)((size_t)(a.Ax1 - bA) * sizeof(R));
    c.rP = f2_rsrc(bP + sb, fb);
    c.rAx = f2_rsrc(bA + sb, fb);
    c.rAz = f2_rsrc(bAz + sb, fb);
    c.rcm = f2_rsrc(a.cm, pb);
    if constexpr (MODE != F2_REC) {
        c.rPx = f2_rsrc((a.Px0 < a.Px1 ? a.Px0 : a.Px1) + sb, fb);
        c.rpm = f2_rsrc(a.pmap, (unsigned)((size_t)a.sstride * 16));
    }
    if constexpr (MODE == F2_FWD)   // ring samples of this shot, time-major (none: no bytes)
        c.rring = f2_rsrc(a.ring ? a.ring + (size_t)c.s * a.T * a.n_ring : a.cm,
                          a.ring ? (unsigned)((size_t)a.T * a.n_ring * sizeof(R)) : 0u);
    if constexpr (MODE != F2_FWD) {   // lap1 = lap0 + slot_elems
        c.rl = f2_rsrc(a.lap0 + sb, fb);
        c.dL1 = (unsigned)((size_t)a.slot_elems * sizeof(R));
    }
    if constexpr (MODE == F2_LAM) c.rg = f2_rsrc(a.grad + sb, pb);
    if constexpr (MODE == F2_REC) {
        // ring samples of this shot, time-major
        c.rring = f2_rsrc(a.ring_in + (size_t)c.s * a.T * a.n_ring,
                          (unsigned)((size_t)a.T * a.n_ring * sizeof(R)));
    }
    if constexpr (MODE != F2_REC)
        for (int r = c.tid; r < F2_NR; r += NW) {
            const int z = c.zc0 - F2_RB + r;
            skz[r] = (z >= -ZOFF && z < a.Z + ZOFF) ? a.kz[ZOFF + z] : R(1);
        }
    if constexpr (MODE != F2_LAM)
        for (int k = c.tid; k < F2_NR * MW; k += NW) {
            const int z = c.zc0 - F2_RB + k / MW;
            smask[k] = (z >= 0 && z < a.Z) ? a.mask[((size_t)id.x() * a.Z + z) * MW + k % MW] : 0u;
        }
    if constexpr (MODE != F2_REC)
        for (int r = c.tid; r < F2_NR; r += NW) {
            const int z = c.zc0 - F2_RB + r;
            sev[r] = (z >= 0 && z < a.Z) ? a.evrow[(size_t)id.x() * a.Z + z] : (uint8_t)0;
        }
    for (int r = 0; r < 8; r++)
        for (int k = c.tid; k < LP; k += NW) {
            sAx0[r][k] = R(0); sP1[r][k] = R(0); sAx1[r][k] = R(0); sP2[r][k] = R(0);
        }
    Pair2Regs<R, MODE> g;
#pragma unroll
    for (int k = 0; k < 8; k++)
        g.az0[k] = g.p1[k] = g.s1[k] = g.az1[k] = g.p2[k] = g.gq[k] = R(0);
#pragma unroll
    for (int k = 0; k < 4; k++) g.p0[k] = g.s0[k] = g.m0[k] = R(0);
    __syncthreads();
    const int i0 = c.zc0 - F2_LEAD;
#pragma unroll
    for (int k = 0; k < f2_pf<MODE>(); k++)
        f2_fetch<R, MODE, NW, false>(c, smask, sev, skz, i0 + k, g.pf[k]);
    const int nit = c.zc1 - c.zc0 + F2_LEAD + 14;
    // Steady blocks of 8 iterations [js, je): iteration j (row i = i0 + j)
    // touches rows i - 14 .. i + f2_pf <= i + 4; all of them lie in the
    // chunk, 4 or more rows from the plane's edges and in the plane when
    // i - 14 >= max(zc0, BOX) and i - 3 < min(zc1, Z - BOX) (then also
    // i + 4 < Z + 4), i.e. j >= max(zc0, BOX) - zc0 + 29 and
    // j <= min(zc1 + 2, Z - 2) - zc0 + 15.
    const int js = (max(c.zc0, BOX) - c.zc0 + 29 + 7) & ~7;
    const int jl = min(c.zc1 + 2, a.Z - 2) - c.zc0 + 15 - 7;   // last start of a steady block
    const int je = jl >= js ? js + ((jl - js) / 8 + 1) * 8 : js;
    for (int j0 = 0; j0 < nit;) {
        if (f2_steady<MODE>() && j0 == js) {
            for (; j0 < je; j0 += 8) {
                const int i = i0 + j0;
                f2_iter<R, MODE, NW, 0, true>(a, c, g, sAx0, sP1, sAx1, sP2, smask, sev, skz, i + 0, kx);
                f2_iter<R, MODE, NW, 1, true>(a, c, g, sAx0, sP1, sAx1, sP2, smask, sev, skz, i + 1, kx);
                f2_iter<R, MODE, NW, 2, true>(a, c, g, sAx0, sP1, sAx1, sP2, smask, sev, skz, i + 2, kx);
                f2_iter<R, MODE, NW, 3, true>(a, c, g, sAx0, sP1, sAx1, sP2, smask, sev, skz, i + 3, kx);
                f2_iter<R, MODE, NW, 4, true>(a, c, g, sAx0, sP1, sAx1, sP2, smask, sev, skz, i + 4, kx);
                f2_iter<R, MODE, NW, 5, true>(a, c, g, sAx0, sP1, sAx1, sP2, smask, sev, skz, i + 5, kx);
                f2_iter<R, MODE, NW, 6, true>(a, c, g, sAx0, sP1, sAx1, sP2, smask, sev, skz, i + 6, kx);
                f2_iter<R, MODE, NW, 7, true>(a, c, g, sAx0, sP1, sAx1, sP2, smask, sev, skz, i + 7, kx);
            }
            if (j0 >= nit) break;
        }
        const int i = i0 + j0;
        f2_iter<R, MODE, NW, 0, false>(a, c, g, sAx0, sP1, sAx1, sP2, smask, sev, skz, i + 0, kx);
        f2_iter<R, MODE, NW, 1, false>(a, c, g, sAx0, sP1, sAx1, sP2, smask, sev, skz, i + 1, kx);
        f2_iter<R, MODE, NW, 2, false>(a, c, g, sAx0, sP1, sAx1, sP2, smask, sev, skz, i + 2, kx);
        f2_iter<R, MODE, NW, 3, false>(a, c, g, sAx0, sP1, sAx1, sP2, smask, sev, skz, i + 3, kx);
        f2_iter<R, MODE, NW, 4, false>(a, c, g, sAx0, sP1, sAx1, sP2, smask, sev, skz, i + 4, kx);
        f2_iter<R, MODE, NW, 5, false>(a, c, g, sAx0, sP1, sAx1, sP2, smask, sev, skz, i + 5, kx);
        f2_iter<R, MODE, NW, 6, false>(a, c, g, sAx0, sP1, sAx1, sP2, smask, sev, skz, i + 6, kx);
        f2_iter<R, MODE, NW, 7, false>(a, c, g, sAx0, sP1, sAx1, sP2, smask, sev, skz, i + 7, kx);
        j0 += 8;
    }
}

template <typename R, int MODE, int NW>
__global__ void __launch_bounds__(NW, (f2_waves<R, MODE, NW>())) pair2_kernel(const Pair2Args<R> a)
{
    pair2_body<R, MODE, NW>(a, F2GridIds{});
}

// One two-step launch over several independent problems (models of uneven
// grid size and shot count, each with its own time step): work-group b runs
// problem i, wg0[i] <= b < wg0[i + 1], at local index l = b - wg0[i] taken as
// (column block, row chunk, shot) with the column block fastest, as the
// grid of a one-problem launch orders them.  The tables are read through the
// constant address space: uniform, so the problem's arguments land in scalar
// registers as a one-problem launch's kernel arguments do.
template <typename R, int MODE, int NW>
__global__ void __launch_bounds__(NW, (f2_waves<R, MODE, NW>()))
pair2_multi_kernel(const F2_CONST Pair2Args<R> *tab, const F2_CONST int *wg0, const int nprob)
{
    const int b = blockIdx.x;
    int i = 0;
    while (i + 1 < nprob && b >= wg0[i + 1]) i++;
    Pair2Args<R> a;
    {
        static_assert(sizeof(Pair2Args<R>) % 4 == 0, "argument block of whole words");
        const F2_CONST unsigned *src = (const F2_CONST unsigned *)(tab + i);
        unsigned *dst = reinterpret_cast<unsigned *>(&a);
#pragma unroll
        for (int k = 0; k < (int)(sizeof(Pair2Args<R>) / 4); k++) dst[k] = src[k];
    }
    const int nbx = (a.X + F2Geo<NW>::OWN - 1) / F2Geo<NW>::OWN;
    const int nby = (a.Z + F2_CH - 1) / F2_CH;
    const unsigned l = (unsigned)(b - wg0[i]);
    pair2_body<R, MODE, NW>(a, F2GivenIds{l % (unsigned)nbx, (l / (unsigned)nbx) % (unsigned)nby,
                                          l / (unsigned)(nbx * nby)});
}

long long pair2_groups(int X, int Z, int nw, int nshots)
{
    const int own = f2_own(nw);
    return own > 0 ? (long long)((X + own - 1) / own) * ((Z + F2_CH - 1) / F2_CH) * nshots : 0;
}

bool pair2_supported(long long sstride, long long slot_elems, int prec)
{
    // buffer offsets below F2_SENT: the event map (16 bytes per cell) and
    // both slots of a field seen from one shot must fit
    return sstride > 0 && sstride * 16 <= (long long)F2_SENT &&
           (slot_elems + sstride) * prec <= (long long)F2_SENT;
}

template <typename R>
bool pair2_args_ok(const Pair2Args<R> &a, int mode)
{
    if (mode != F2_FWD && mode != F2_LAM && mode != F2_REC) return false;
    if ((mode != F2_LAM && !a.mask) || (mode != F2_REC && (!a.pmap || !a.evrow)) ||
        !pair2_supported(a.sstride, a.slot_elems, (int)sizeof(R)))
        return false;
    if (a.X <= 0 || a.Z <= 0 || a.nshots <= 0) return false;
    // both slots of each field one allocation: the written slot is slot_elems away
    auto gap = [&](const R *u, const R *v) { return u > v ? u - v : v - u; };
    if (gap(a.P0, a.P1) != a.slot_elems || gap(a.Ax0, a.Ax1) != a.slot_elems ||
        gap(a.Az0, a.Az1) != a.slot_elems || (a.Ax0 < a.Ax1) != (a.Az0 < a.Az1) ||
        (mode != F2_FWD && a.lap1 - a.lap0 != a.slot_elems) ||
        (mode != F2_REC && (gap(a.Px0, a.Px1) != a.slot_elems || (a.P0 < a.P1) != (a.Px0 < a.Px1))))
        return false;
    if (mode != F2_LAM && (size_t)a.T * a.n_ring * sizeof(R) > F2_SENT) return false;
    return f2_own(a.nw) > 0;
}
template bool pair2_args_ok<float>(const Pair2Args<float> &, int);
template bool pair2_args_ok<double>(const Pair2Args<double> &, int);

template <typename R>
hipError_t launch_pair2(const Pair2Args<R> &a, int mode, hipStream_t st)
{
    if (!pair2_args_ok(a, mode)) return hipErrorInvalidValue;
    const int own = f2_own(a.nw);
    dim3 grid((a.X + own - 1) / own, (a.Z + F2_CH - 1) / F2_CH, a.nshots);
    const dim3 blk(a.nw);
#define F2_LAUNCH(M, W) hipLaunchKernelGGL((pair2_kernel<R, M, W>), grid, blk, 0, st, a)
#define F2_WIDTHS(M)                                      \
    switch (a.nw) {                                       \
    case 128: F2_LAUNCH(M, 128); break;                   \
    case 192: F2_LAUNCH(M, 192); break;                   \
    case 256: F2_LAUNCH(M, 256); break;                   \
    case 384: F2_LAUNCH(M, 384); break;                   \
    default: return hipErrorInvalidValue;                 \
    }
    switch (mode) {
    case F2_FWD: F2_WIDTHS(F2_FWD); break;
    case F2_LAM: F2_WIDTHS(F2_LAM); break;
    case F2_REC: F2_WIDTHS(F2_REC); break;
    default: return hipErrorInvalidValue;
    }
#undef F2_WIDTHS
#undef F2_LAUNCH
    return hipGetLastError();
}
template hipError_t launch_pair2<float>(const Pair2Args<float> &, int, hipStream_t);
template hipError_t launch_pair2<double>(const Pair2Args<double> &, int, hipStream_t);

hipError_t launch_pair2_multi(const Pair2Args<float> *tab, const int *wg0, int nprob, long long nwg,
                              int mode, hipStream_t st)
{
    if (!tab || !wg0 || nprob <= 0 || nwg <= 0 || nwg > 0x7FFFFFFFLL) return hipErrorInvalidValue;
    const dim3 grid((unsigned)nwg), blk(F2_MULTI_NW);
    const F2_CONST Pair2Args<float> *ct = (const F2_CONST Pair2Args<float> *)tab;
    const F2_CONST int *cw = (const F2_CONST int *)wg0;
    switch (mode) {
    case F2_FWD:
        hipLaunchKernelGGL((pair2_multi_kernel<float, F2_FWD, F2_MULTI_NW>), grid, blk, 0, st, ct, cw, nprob);
        break;
    case F2_LAM:
        hipLaunchKernelGGL((pair2_multi_kernel<float, F2_LAM, F2_MULTI_NW>), grid, blk, 0, st, ct, cw, nprob);
        break;
    case F2_REC:
        hipLaunchKernelGGL((pair2_multi_kernel<float, F2_REC, F2_MULTI_NW>), grid, blk, 0, st, ct, cw, nprob);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int step_occupancy(bool adj, int prec, int th, bool persistent)
{
    if (prec == 8) {
        if (!adj && th == 16) return occ_fwd<double, 16>(persistent);
        if (!adj && th == 32) return occ_fwd<double, 32>(persistent);
        if (adj && th == 16) return occ_adj<double, 16>(persistent);
        return 0;
    }
    switch (th) {
    case 16: return adj ? occ_adj<float, 16>(persistent) : occ_fwd<float, 16>(persistent);
    case 32: return adj ? occ_adj<float, 32>(persistent) : occ_fwd<float, 32>(persistent);
    case 64: return adj ? occ_adj<float, 64>(persistent) : occ_fwd<float, 64>(persistent);
    default: return 0;
    }
}

}  // namespace otfwi
