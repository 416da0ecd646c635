// Fused time-step kernels of the split-field PML acoustic propagator.
//
// Scheme (reference: simulator/src/regression_PML_single.cu:447-615 with the
// prec_deriv==4 coefficients of simulator/src/deriv_macros.h:84-108):
//   P half-step  Px(t) = Px(t-1)*kx + m*D-x Ax(t-1),  Pz(t) = Pz(t-1)*kz + m*D-z Az(t-1)
//   A half-step  Ax(t) = Ax(t-1)*kx + cA*D+x P(t),    Az(t) = Az(t-1)*kz + cA*D+z P(t)
// with P = Px+Pz, m = (c^2 dt)/dx, cA = dt/dx, k = 1 - d*dt looked up one
// cell further along its axis, and every stencil operand outside the box
// 4 <= x < X-4, 4 <= z < Z-4 reading 0.
//
// Storage: P = Px+Pz is kept on every cell; the split component Px (the
// boundary memory variable "psi") is read and written only on damped cells,
// where kx != 1 or kz != 1.  Ax and Az are full fields.  Each field is
// [slot][shot][Z][ld] FP32 (or FP64 for the validation build), x fastest.
//
// One launch advances every shot of a batch by one full time step.  A
// work-group owns a 64 x TH tile: it stages A(t-1) with an 8-cell halo in
// LDS, recomputes P(t) on the tile plus a 4-cell halo (the "E box"), keeps
// that in LDS and finishes A(t) on the tile, so the per-step HBM traffic is
// the algorithmic minimum (read P, Ax, Az, m; write P, Ax, Az) plus halo
// re-reads that hit in L2.  Source injection, receiver sampling and the
// boundary-ring save happen on the cells where they occur, inside the same
// kernel.
#include "otfwi_internal.h"

namespace otfwi {

// Coefficients of the 8th-order staggered first derivative exactly as the
// reference writes them (double literals, deriv_macros.h:87-94).
constexpr double K1 = 1.196289e+00, K2 = -7.975260e-02, K3 = 9.570313e-03, K4 = -6.975446e-04;

// sum_k c_k (v[3+k] - v[4-k]) for taps v[j] = f(x0 + j): the backward
// operator D- at x = x0+4 and the forward operator D+ at x = x0+3.
//
// FP32: each coefficient is carried as hi + lo FP32 parts.  Rounding the
// coefficients to FP32 alone changes the stencil systematically, which moves
// the adjoint gradient by ~1e-4 relative over a 1500-step run; with the lo
// parts folded in first and the hi terms accumulated smallest first, the
// FP32 derivative is as accurate as the reference's FP64-evaluated one.
template <typename R>
__device__ __forceinline__ R stag(R v0, R v1, R v2, R v3, R v4, R v5, R v6, R v7);

template <>
__device__ __forceinline__ float stag<float>(float v0, float v1, float v2, float v3, float v4,
                                             float v5, float v6, float v7)
{
    constexpr float h1 = (float)K1, h2 = (float)K2, h3 = (float)K3, h4 = (float)K4;
    constexpr float l1 = (float)(K1 - (double)h1), l2 = (float)(K2 - (double)h2);
    constexpr float l3 = (float)(K3 - (double)h3), l4 = (float)(K4 - (double)h4);
    const float d1 = v4 - v3, d2 = v5 - v2, d3 = v6 - v1, d4 = v7 - v0;
    float s = l1 * d1;
    s = fmaf(l2, d2, s);
    s = fmaf(l3, d3, s);
    s = fmaf(l4, d4, s);
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

__device__ __forceinline__ bool in_box(int x, int z, int X, int Z)
{
    return x >= BOX && z >= BOX && x < X - BOX && z < Z - BOX;
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

#define ROW8(A, r, c) (A)[r][(c) + 0], (A)[r][(c) + 1], (A)[r][(c) + 2], (A)[r][(c) + 3], \
                      (A)[r][(c) + 4], (A)[r][(c) + 5], (A)[r][(c) + 6], (A)[r][(c) + 7]

template <typename R, int TH>
__global__ void __launch_bounds__(NTHREADS)
fwd_step_kernel(const FwdArgs<R> a)
{
    constexpr int EH = TH + 8, AZH = TH + 16, RE = EH / 4, RO = TH / 4;
    static_assert(TH % 4 == 0, "tile height must be a multiple of 4");
    __shared__ R sAx[EH][AXW];   // Ax(t-1): rows z0-4.., cols x0-8..
    __shared__ R sAz[AZH][EW];   // Az(t-1): rows z0-8.., cols x0-4..
    __shared__ R sP[EH][EW];     // P(t) on the E box: rows z0-4.., cols x0-4..

    const int tiles = a.ntx * a.ntz;
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int s = L / tiles, tl = L - s * tiles;
    const int tz = tl / a.ntx, tx = tl - tz * a.ntx;
    const int x0 = tx * TW, z0 = tz * TH;
    const int X = a.X, Z = a.Z, ld = a.ld;
    const size_t so = (size_t)s * (size_t)a.sstride;
    const int tid = threadIdx.x;
    const uint8_t fl = a.tflags[tl];

    {
        const R *Ax0 = a.Ax0 + so, *Az0 = a.Az0 + so;
        for (int i = tid; i < EH * AXW; i += NTHREADS) {
            const int r = i / AXW, c = i - r * AXW;
            const int z = z0 - 4 + r, x = x0 - 8 + c;
            sAx[r][c] = in_box(x, z, X, Z) ? Ax0[z * ld + x] : R(0);
        }
        for (int i = tid; i < AZH * EW; i += NTHREADS) {
            const int r = i / EW, c = i - r * EW;
            const int z = z0 - 8 + r, x = x0 - 4 + c;
            sAz[r][c] = in_box(x, z, X, Z) ? Az0[z * ld + x] : R(0);
        }
    }

    // ---- E phase: P(t) on the tile plus halo -------------------------------
    const bool eact = tid < EW * 4;
    const int ex = tid % EW, ey = tid / EW;
    const int xe = x0 - 4 + ex;
    const bool xin = xe >= BOX && xe < X - BOX;
    const int er0 = ey * RE;
    R pp[RE], mm[RE];
    if (eact) {
#pragma unroll
        for (int j = 0; j < RE; j++) {
            const int z = z0 - 4 + er0 + j;
            const bool in = xin && z >= BOX && z < Z - BOX;
            pp[j] = in ? a.P0[so + z * ld + xe] : R(0);
            mm[j] = in ? a.cm[z * ld + xe] : R(0);
        }
    }
    __syncthreads();

    if (eact) {
        const R kxv = ((fl & TF_STRIP) && xin) ? a.kx[xe] : R(1);
        R w0 = sAz[er0 + 0][ex], w1 = sAz[er0 + 1][ex], w2 = sAz[er0 + 2][ex], w3 = sAz[er0 + 3][ex];
        R w4 = sAz[er0 + 4][ex], w5 = sAz[er0 + 5][ex], w6 = sAz[er0 + 6][ex], w7 = sAz[er0 + 7][ex];
#pragma unroll
        for (int j = 0; j < RE; j++) {
            const int er = er0 + j;
            const int z = z0 - 4 + er;
            const bool in = xin && z >= BOX && z < Z - BOX;
            const R dAx = stag<R>(ROW8(sAx, er, ex));
            const R dAz = stag<R>(w0, w1, w2, w3, w4, w5, w6, w7);
            w0 = w1; w1 = w2; w2 = w3; w3 = w4; w4 = w5; w5 = w6; w6 = w7;
            w7 = sAz[er + 8][ex];
            const int gi = z * ld + xe;
            R psi = R(0), kzv = R(1);
            bool strip = false;
            if (fl & TF_STRIP) {
                if (in) kzv = a.kz[z];
                strip = in && (kxv != R(1) || kzv != R(1));
                if (strip) psi = a.Px0[so + gi];
            }
            R nx = fma(psi, kxv, mm[j] * dAx);
            R nz = fma(pp[j] - psi, kzv, mm[j] * dAz);
            const bool owned = er >= 4 && er < TH + 4 && ex >= 4 && ex < TW + 4 && xe < X && z < Z;
            if ((fl & TF_EVENT) && in) {
                const int e = a.evt[gi];
                if (e >= 0) {
                    const PointEvt pe = a.pts[e];
                    if (pe.src >= 0) {
                        const int ss = a.shot_src[s];
                        if (ss < 0 || ss == pe.src) {
                            const R h = a.icum[(size_t)pe.src * a.T + a.t] * R(0.5);
                            nx += h;
                            nz += h;
                        }
                    }
                    if (owned) {
                        const R pv = nx + nz;
                        if (pe.rx >= 0 && a.rec)
                            a.rec[((size_t)s * a.n_rx + pe.rx) * a.T + a.t] = pv;
                        if (pe.ring >= 0 && a.ring)
                            a.ring[((size_t)s * a.T + a.t) * a.n_ring + pe.ring] = pv;
                    }
                }
            }
            const R p = in ? nx + nz : R(0);
            sP[er][ex] = p;
            if (owned) {
                a.P1[so + gi] = p;
                if (strip) a.Px1[so + gi] = nx;
            }
        }
    }
    __syncthreads();

    // ---- owned phase: A(t) on the tile ------------------------------------
    if (tid < TW * 4) {
        const int ox = tid & (TW - 1), oy = tid / TW;
        const int x = x0 + ox;
        const bool xin2 = x >= BOX && x < X - BOX;
        const R kxv = ((fl & TF_STRIP) && xin2) ? a.kx[x] : R(1);
        const int or0 = oy * RO;
        R w0 = sP[or0 + 1][ox + 4], w1 = sP[or0 + 2][ox + 4], w2 = sP[or0 + 3][ox + 4], w3 = sP[or0 + 4][ox + 4];
        R w4 = sP[or0 + 5][ox + 4], w5 = sP[or0 + 6][ox + 4], w6 = sP[or0 + 7][ox + 4], w7 = sP[or0 + 8][ox + 4];
#pragma unroll
        for (int j = 0; j < RO; j++) {
            const int orr = or0 + j;
            const int z = z0 + orr;
            const R dPx = stag<R>(ROW8(sP, orr + 4, ox + 1));
            const R dPz = stag<R>(w0, w1, w2, w3, w4, w5, w6, w7);
            if (j + 1 < RO) {
                w0 = w1; w1 = w2; w2 = w3; w3 = w4; w4 = w5; w5 = w6; w6 = w7;
                w7 = sP[orr + 9][ox + 4];
            }
            if (x < X && z < Z) {
                const bool in = xin2 && z >= BOX && z < Z - BOX;
                const R kzv = ((fl & TF_STRIP) && in) ? a.kz[z] : R(1);
                const int gi = z * ld + x;
                const R ax = sAx[orr + 4][ox + 8], az = sAz[orr + 8][ox + 4];
                a.Ax1[so + gi] = in ? fma(ax, kxv, a.cA * dPx) : R(0);
                a.Az1[so + gi] = in ? fma(az, kzv, a.cA * dPz) : R(0);
            }
        }
    }
}

// Adjoint step (reference phase 2, regression_PML_single.cu:793-823): the
// adjoint field lambda advances with the damped forward operator and takes the
// running sum of the time-flipped adjoint source at the receivers; the forward
// field is reconstructed backwards with the undamped reverse operator and the
// saved boundary ring; grad += lambda_P(t) * lap8(U(t)) on the tile.
template <typename R, int TH>
__global__ void __launch_bounds__(NTHREADS)
adj_step_kernel(const AdjArgs<R> a)
{
    constexpr int EH = TH + 8, AZH = TH + 16, RE = EH / 4, RO = TH / 4;
    __shared__ R sLAx[EH][AXW];
    __shared__ R sLAz[AZH][EW];
    __shared__ R sLP[EH][EW];
    __shared__ R sUAx[EH][AXW];
    __shared__ R sUAz[AZH][EW];
    __shared__ R sU[EH][EW];

    const int tiles = a.ntx * a.ntz;
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int s = L / tiles, tl = L - s * tiles;
    const int tz = tl / a.ntx, tx = tl - tz * a.ntx;
    const int x0 = tx * TW, z0 = tz * TH;
    const int X = a.X, Z = a.Z, ld = a.ld;
    const size_t so = (size_t)s * (size_t)a.sstride;
    const int tid = threadIdx.x;
    const uint8_t fl = a.tflags[tl];
    const int trev = a.T - 1 - a.t;   // forward time of U(t)

    {
        const R *LAx0 = a.LAx0 + so, *LAz0 = a.LAz0 + so;
        const R *UAx0 = a.UAx0 + so, *UAz0 = a.UAz0 + so;
        for (int i = tid; i < EH * AXW; i += NTHREADS) {
            const int r = i / AXW, c = i - r * AXW;
            const int z = z0 - 4 + r, x = x0 - 8 + c;
            const bool in = in_box(x, z, X, Z);
            sLAx[r][c] = in ? LAx0[z * ld + x] : R(0);
            sUAx[r][c] = in ? UAx0[z * ld + x] : R(0);
        }
        for (int i = tid; i < AZH * EW; i += NTHREADS) {
            const int r = i / EW, c = i - r * EW;
            const int z = z0 - 8 + r, x = x0 - 4 + c;
            const bool in = in_box(x, z, X, Z);
            sLAz[r][c] = in ? LAz0[z * ld + x] : R(0);
            sUAz[r][c] = in ? UAz0[z * ld + x] : R(0);
        }
    }

    const bool eact = tid < EW * 4;
    const int ex = tid % EW, ey = tid / EW;
    const int xe = x0 - 4 + ex;
    const bool xin = xe >= BOX && xe < X - BOX;
    const int er0 = ey * RE;
    R lp[RE], up[RE], mm[RE];
    if (eact) {
#pragma unroll
        for (int j = 0; j < RE; j++) {
            const int z = z0 - 4 + er0 + j;
            const bool in = xin && z >= BOX && z < Z - BOX;
            lp[j] = in ? a.LP0[so + z * ld + xe] : R(0);
            up[j] = in ? a.U0[so + z * ld + xe] : R(0);
            mm[j] = in ? a.cm[z * ld + xe] : R(0);
        }
    }
    __syncthreads();

    if (eact) {
        const R kxv = ((fl & TF_STRIP) && xin) ? a.kx[xe] : R(1);
        R w0 = sLAz[er0 + 0][ex], w1 = sLAz[er0 + 1][ex], w2 = sLAz[er0 + 2][ex], w3 = sLAz[er0 + 3][ex];
        R w4 = sLAz[er0 + 4][ex], w5 = sLAz[er0 + 5][ex], w6 = sLAz[er0 + 6][ex], w7 = sLAz[er0 + 7][ex];
        R v0 = sUAz[er0 + 0][ex], v1 = sUAz[er0 + 1][ex], v2 = sUAz[er0 + 2][ex], v3 = sUAz[er0 + 3][ex];
        R v4 = sUAz[er0 + 4][ex], v5 = sUAz[er0 + 5][ex], v6 = sUAz[er0 + 6][ex], v7 = sUAz[er0 + 7][ex];
#pragma unroll
        for (int j = 0; j < RE; j++) {
            const int er = er0 + j;
            const int z = z0 - 4 + er;
            const bool in = xin && z >= BOX && z < Z - BOX;
            const R dLx = stag<R>(ROW8(sLAx, er, ex));
            const R dLz = stag<R>(w0, w1, w2, w3, w4, w5, w6, w7);
            const R dUx = stag<R>(ROW8(sUAx, er, ex));
            const R dUz = stag<R>(v0, v1, v2, v3, v4, v5, v6, v7);
            w0 = w1; w1 = w2; w2 = w3; w3 = w4; w4 = w5; w5 = w6; w6 = w7;
            w7 = sLAz[er + 8][ex];
            v0 = v1; v1 = v2; v2 = v3; v3 = v4; v4 = v5; v5 = v6; v6 = v7;
            v7 = sUAz[er + 8][ex];
            const int gi = z * ld + xe;
            R psi = R(0), kzv = R(1);
            bool strip = false;
            if (fl & TF_STRIP) {
                if (in) kzv = a.kz[z];
                strip = in && (kxv != R(1) || kzv != R(1));
                if (strip) psi = a.LPx0[so + gi];
            }
            R nx = fma(psi, kxv, mm[j] * dLx);
            R nz = fma(lp[j] - psi, kzv, mm[j] * dLz);
            R u = fma(-mm[j], dUx + dUz, up[j]);
            const bool owned = er >= 4 && er < TH + 4 && ex >= 4 && ex < TW + 4 && xe < X && z < Z;
            if ((fl & TF_EVENT) && in) {
                const int e = a.evt[gi];
                if (e >= 0) {
                    const PointEvt pe = a.pts[e];
                    if (pe.rx >= 0) {
                        const R h = a.jcum[((size_t)s * a.n_rx + pe.rx) * a.T + a.t] * R(0.5);
                        nx += h;
                        nz += h;
                    }
                    if (pe.ring >= 0)
                        u = a.ring[((size_t)s * a.T + trev) * a.n_ring + pe.ring];
                }
            }
            const R lpv = in ? nx + nz : R(0);
            const R uv = in ? u : R(0);
            sLP[er][ex] = lpv;
            sU[er][ex] = uv;
            if (owned) {
                a.LP1[so + gi] = lpv;
                a.U1[so + gi] = uv;
                if (strip) a.LPx1[so + gi] = nx;
            }
        }
    }
    __syncthreads();

    if (tid < TW * 4) {
        const int ox = tid & (TW - 1), oy = tid / TW;
        const int x = x0 + ox;
        const bool xin2 = x >= BOX && x < X - BOX;
        const R kxv = ((fl & TF_STRIP) && xin2) ? a.kx[x] : R(1);
        const int or0 = oy * RO;
        R w0 = sLP[or0 + 1][ox + 4], w1 = sLP[or0 + 2][ox + 4], w2 = sLP[or0 + 3][ox + 4], w3 = sLP[or0 + 4][ox + 4];
        R w4 = sLP[or0 + 5][ox + 4], w5 = sLP[or0 + 6][ox + 4], w6 = sLP[or0 + 7][ox + 4], w7 = sLP[or0 + 8][ox + 4];
        // U window covers rows orr .. orr+8 (the 9-point Laplacian); the
        // staggered derivative uses the upper eight.
        R u0 = sU[or0 + 0][ox + 4], u1 = sU[or0 + 1][ox + 4], u2 = sU[or0 + 2][ox + 4], u3 = sU[or0 + 3][ox + 4];
        R u4 = sU[or0 + 4][ox + 4], u5 = sU[or0 + 5][ox + 4], u6 = sU[or0 + 6][ox + 4], u7 = sU[or0 + 7][ox + 4];
        R u8 = sU[or0 + 8][ox + 4];
        R *G = a.grad + so;
#pragma unroll
        for (int j = 0; j < RO; j++) {
            const int orr = or0 + j;
            const int z = z0 + orr;
            const R dLPx = stag<R>(ROW8(sLP, orr + 4, ox + 1));
            const R dLPz = stag<R>(w0, w1, w2, w3, w4, w5, w6, w7);
            const R dUx = stag<R>(ROW8(sU, orr + 4, ox + 1));
            const R dUz = stag<R>(u1, u2, u3, u4, u5, u6, u7, u8);
            const R lx = lap9<R>(sU[orr + 4][ox + 0], sU[orr + 4][ox + 1], sU[orr + 4][ox + 2],
                                 sU[orr + 4][ox + 3], u4, sU[orr + 4][ox + 5], sU[orr + 4][ox + 6],
                                 sU[orr + 4][ox + 7], sU[orr + 4][ox + 8]);
            const R lz = lap9<R>(u0, u1, u2, u3, u4, u5, u6, u7, u8);
            const R lam = sLP[orr + 4][ox + 4];
            if (j + 1 < RO) {
                w0 = w1; w1 = w2; w2 = w3; w3 = w4; w4 = w5; w5 = w6; w6 = w7;
                w7 = sLP[orr + 9][ox + 4];
                u0 = u1; u1 = u2; u2 = u3; u3 = u4; u4 = u5; u5 = u6; u6 = u7; u7 = u8;
                u8 = sU[orr + 9][ox + 4];
            }
            if (x < X && z < Z) {
                const bool in = xin2 && z >= BOX && z < Z - BOX;
                const R kzv = ((fl & TF_STRIP) && in) ? a.kz[z] : R(1);
                const int gi = z * ld + x;
                const R lax = sLAx[orr + 4][ox + 8], laz = sLAz[orr + 8][ox + 4];
                const R uax = sUAx[orr + 4][ox + 8], uaz = sUAz[orr + 8][ox + 4];
                a.LAx1[so + gi] = in ? fma(lax, kxv, a.cA * dLPx) : R(0);
                a.LAz1[so + gi] = in ? fma(laz, kzv, a.cA * dLPz) : R(0);
                a.UAx1[so + gi] = in ? fma(-a.cA, dUx, uax) : R(0);
                a.UAz1[so + gi] = in ? fma(-a.cA, dUz, uaz) : R(0);
                G[gi] = fma(lam, (lx + lz) * (R(1) / R(5040)), G[gi]);
            }
        }
    }
}

template <typename R, int TH>
static hipError_t fwd_launch_th(const FwdArgs<R> &a, hipStream_t st)
{
    const int nb = a.ntx * a.ntz * a.nshots;
    hipLaunchKernelGGL((fwd_step_kernel<R, TH>), dim3(nb), dim3(NTHREADS), 0, st, a);
    return hipGetLastError();
}

template <typename R, int TH>
static hipError_t adj_launch_th(const AdjArgs<R> &a, hipStream_t st)
{
    const int nb = a.ntx * a.ntz * a.nshots;
    hipLaunchKernelGGL((adj_step_kernel<R, TH>), dim3(nb), dim3(NTHREADS), 0, st, a);
    return hipGetLastError();
}

template <>
hipError_t launch_fwd_step<float>(const FwdArgs<float> &a, int th, hipStream_t s)
{
    switch (th) {
    case 16: return fwd_launch_th<float, 16>(a, s);
    case 32: return fwd_launch_th<float, 32>(a, s);
    case 64: return fwd_launch_th<float, 64>(a, s);
    default: return hipErrorInvalidValue;
    }
}
template <>
hipError_t launch_fwd_step<double>(const FwdArgs<double> &a, int th, hipStream_t s)
{
    switch (th) {
    case 16: return fwd_launch_th<double, 16>(a, s);
    case 32: return fwd_launch_th<double, 32>(a, s);
    default: return hipErrorInvalidValue;
    }
}
template <>
hipError_t launch_adj_step<float>(const AdjArgs<float> &a, int th, hipStream_t s)
{
    switch (th) {
    case 16: return adj_launch_th<float, 16>(a, s);
    case 32: return adj_launch_th<float, 32>(a, s);
    case 64: return adj_launch_th<float, 64>(a, s);
    default: return hipErrorInvalidValue;
    }
}
template <>
hipError_t launch_adj_step<double>(const AdjArgs<double> &a, int th, hipStream_t s)
{
    switch (th) {
    case 16: return adj_launch_th<double, 16>(a, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace otfwi
