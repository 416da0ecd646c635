// Second-order C-PML pressure scheme (simulator/src/regression_CPML.cu,
// prec_deriv = 4) for gfx950: libotfwi_cpml.so.  Forward only.
//
// Per step t >= 2, on the box 4 <= x < X-4, 4 <= z < Z-4 (values outside
// the box read as 0):
//   frame 1 (simulateFrame1, :46-120): vx0 = (float)D+x(P(t-1)) / dx,
//     mx(t) = bxh mx(t-1) + axh vx0, vx1 = (vx0 / Kxh + mx(t)) / rho_half_x,
//     and the same along z;
//   frame 2 (simulateFrame2, :123-210): vxx0 = (float)D-x(vx1) / dx,
//     mxx(t) = bx mxx(t-1) + ax vxx0, vxx1 = vxx0 / Kx + mxx(t), likewise z,
//     P(t) = 2 P(t-1) - P(t-2) + (vxx1 + vzz1) dt2 kappa;
//   sources (somaFonte, :213-237): P(t)[src] += qt_pi cquad[src] s[t] dt2;
// and receivers record P(t) for every t (gravaBufferSensores2ordem, :240-248).
// D+ / D- are the 8-point staggered differences of deriv_macros.h:85-108
// (offset 0 / 1), double literals times floats summed in double, then
// rounded to float, as the reference evaluates them.
//
// A launch advances every shot of a batch (grid z); each shot has its own
// fields.  P keeps four time slots as the reference does (t % 4; the first
// two from `initial`), the memory variables two (a step reads only t - 1).
// One launch per step (both frames) on 64 x 32-cell work-group tiles whose
// stencil inputs and derivative fields live in LDS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "otfwi_cpml.h"

namespace {

thread_local std::string g_err;

int cfail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}
#define CHK(expr)                                                                      \
    do {                                                                               \
        const hipError_t e_ = (expr);                                                  \
        if (e_ != hipSuccess) return cfail(-2, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

constexpr double C1 = 1.196289e+00, C2 = 7.975260e-02, C3 = 9.570313e-03, C4 = 6.975446e-04;

struct Dims {
    int X, Z;
    long long plane;   // X * Z
};

__device__ __forceinline__ bool inbox(const Dims &d, int x, int z)
{
    return x >= 4 && z >= 4 && x < d.X - 4 && z < d.Z - 4;
}
// f at (x, z) of one plane, 0 outside the box
__device__ __forceinline__ float boxv(const float *f, const Dims &d, int x, int z)
{
    return inbox(d, x, z) ? f[(long long)z * d.X + x] : 0.0f;
}
struct Fields {
    float *P;          // [S][4][plane]
    float *mx, *mz;    // [S][2][plane]
    float *mxx, *mzz;  // [S][2][plane]
};

struct Medium {
    const float *cq, *kappa, *rho_hx, *rho_hz;           // [plane]
    const float *Kx, *ax, *bx, *Kxh, *axh, *bxh;         // [X]
    const float *Kz, *az, *bz, *Kzh, *azh, *bzh;         // [Z]
    float dx, dz, dt2, qt_pi;
};

// derivPMLdx / derivPMLdz (deriv_macros.h:85-108) of 8 staged values, the
// reference's term order: v[0..8) = taps -3-off .. 4-off
__device__ __forceinline__ float dsum(const float *v)
{
    return (float)(-C4 * (double)v[7] + C3 * (double)v[6] - C2 * (double)v[5] +
                   C1 * (double)v[4] - C1 * (double)v[3] + C2 * (double)v[2] -
                   C3 * (double)v[1] + C4 * (double)v[0]);
}

// A memory variable whose profile has a = 0 and b = 1 starts at 0 and stays
// exactly 0 (b m + a v = m): it is neither read nor written there.
__device__ __forceinline__ bool live(float a, float b) { return a != 0.0f || b != 1.0f; }

// One time step t >= 2 of every shot (grid z): frame 1 and frame 2 fused.
// A work-group of 256 threads owns a FX x FZ tile.  It stages P(t-1) with a
// 7-cell halo in LDS (box rule applied while loading), computes the x
// derivative field vx1 on the tile's rows over columns x0-4 .. x0+FX+2 and the
// z field vz1 on its columns over rows z0-4 .. z0+FZ+2 into registers (the
// halo cells recomputed from their neighbours' P, memory variables read but
// only the owner's written), then stores them in LDS over the staged P and
// runs frame 2 on the owned cells.  vx1 and vz1 never reach memory; sharing
// the LDS keeps 19 KB per work-group (8 per CU).  Neighbouring tiles read
// slot t-1 of P and of the memory variables while owners write slot t, so one
// launch per step suffices.
#ifndef CPML_FX
#define CPML_FX 64
#endif
#ifndef CPML_FZ
#define CPML_FZ 32
#endif
constexpr int FX = CPML_FX, FZ = CPML_FZ, FH = FZ * FX / 256;   // FH rows per thread in frame 2
static_assert(256 % FX == 0 && (FZ * FX) % 256 == 0, "tile must split over 256 threads");
constexpr int PW = FX + 14, PH = FZ + 14;   // staged P(t-1): x0-7 .., z0-7 ..
constexpr int VXW = FX + 7, VZH = FZ + 7;   // vx1 columns, vz1 rows in use
constexpr int SXW = FX + 8;                 // row pitch of the vx1 block
constexpr int NVX = (FZ * VXW + 255) / 256, NVZ = (VZH * FX + 255) / 256;
constexpr int LDS_P = PH * PW, LDS_V = FZ * SXW + (FZ + 8) * FX;
constexpr int LDS_N = LDS_P > LDS_V ? LDS_P : LDS_V;

__global__ void __launch_bounds__(256, 8)   // 8 waves per SIMD: at most 64 VGPRs
cpml_step(Fields f, Medium m, Dims d, int t)
{
    __shared__ float lds[LDS_N];
    float (*sp)[PW] = reinterpret_cast<float (*)[PW]>(lds);
    float (*sx)[SXW] = reinterpret_cast<float (*)[SXW]>(lds);              // vx1 at (x0 - 4 + c, z0 + r)
    float (*sz)[FX] = reinterpret_cast<float (*)[FX]>(lds + FZ * SXW);     // vz1 at (x0 + c, z0 - 4 + r)
    const int tid = threadIdx.x, s = blockIdx.z;
    const int x0 = blockIdx.x * FX, z0 = blockIdx.y * FZ;
    const long long pl = d.plane;
    const float *P1 = f.P + ((long long)s * 4 + (t - 1) % 4) * pl;
    for (int k = tid; k < PH * PW; k += 256) {
        const int r = k / PW, c = k - r * PW;
        const bool used = (r >= 7 && r < 7 + FZ) || (c >= 7 && c < 7 + FX);   // not a corner
        sp[r][c] = used ? boxv(P1, d, x0 - 7 + c, z0 - 7 + r) : 0.0f;
    }
    __syncthreads();
    const int c1 = t & 1, p1 = c1 ^ 1;
    const float *mxp = f.mx + ((long long)s * 2 + p1) * pl, *mzp = f.mz + ((long long)s * 2 + p1) * pl;
    float *mxc = f.mx + ((long long)s * 2 + c1) * pl, *mzc = f.mz + ((long long)s * 2 + c1) * pl;
    // frame 1, x: vx1 = (D+x P / dx / Kxh + mx(t)) / rho_hx, taps x-3 .. x+4
    float vx[NVX], vz[NVZ];
#pragma unroll
    for (int n = 0; n < NVX; n++) {
        const int k = tid + n * 256;
        const int r = k / VXW, c = k - r * VXW;
        const int x = x0 - 4 + c, z = z0 + r;
        float out = 0.0f;
        if (k < FZ * VXW && inbox(d, x, z)) {
            const long long i = (long long)z * d.X + x;
            float v[8];
#pragma unroll
            for (int q = 0; q < 8; q++) v[q] = sp[r + 7][c + q];
            const float v0 = dsum(v) / m.dx;
            float mt = 0.0f;
            if (live(m.axh[x], m.bxh[x])) {
                mt = m.bxh[x] * mxp[i] + m.axh[x] * v0;
                if (c >= 4 && c < 4 + FX) mxc[i] = mt;
            }
            out = (v0 / m.Kxh[x] + mt) / m.rho_hx[i];
        }
        vx[n] = out;
    }
    // frame 1, z: vz1 = (D+z P / dz / Kzh + mz(t)) / rho_hz, taps z-3 .. z+4
#pragma unroll
    for (int n = 0; n < NVZ; n++) {
        const int k = tid + n * 256;
        const int r = k / FX, c = k - r * FX;
        const int x = x0 + c, z = z0 - 4 + r;
        float out = 0.0f;
        if (k < VZH * FX && inbox(d, x, z)) {
            const long long i = (long long)z * d.X + x;
            float w[8];
#pragma unroll
            for (int q = 0; q < 8; q++) w[q] = sp[r + q][c + 7];
            const float w0 = dsum(w) / m.dz;
            float mt = 0.0f;
            if (live(m.azh[z], m.bzh[z])) {
                mt = m.bzh[z] * mzp[i] + m.azh[z] * w0;
                if (r >= 4 && r < 4 + FZ) mzc[i] = mt;
            }
            out = (w0 / m.Kzh[z] + mt) / m.rho_hz[i];
        }
        vz[n] = out;
    }
    const int tx = tid & (FX - 1), zg = (tid / FX) * FH;
    float pm1[FH];
#pragma unroll
    for (int j = 0; j < FH; j++) pm1[j] = sp[zg + j + 7][tx + 7];
    __syncthreads();   // every read of the staged P is done
#pragma unroll
    for (int n = 0; n < NVX; n++) {
        const int k = tid + n * 256;
        const int r = k / VXW, c = k - r * VXW;
        if (k < FZ * VXW) sx[r][c] = vx[n];
    }
#pragma unroll
    for (int n = 0; n < NVZ; n++) {
        const int k = tid + n * 256;
        const int r = k / FX, c = k - r * FX;
        if (k < VZH * FX) sz[r][c] = vz[n];
    }
    __syncthreads();
    // frame 2 on the owned cells: taps x-4 .. x+3 of vx1, z-4 .. z+3 of vz1
    const int x = x0 + tx;
    float *mxxp = f.mxx + ((long long)s * 2 + p1) * pl, *mzzp = f.mzz + ((long long)s * 2 + p1) * pl;
    float *mxxc = f.mxx + ((long long)s * 2 + c1) * pl, *mzzc = f.mzz + ((long long)s * 2 + c1) * pl;
    float *P = f.P + (long long)s * 4 * pl;
    const bool kx = x < d.X && live(m.ax[x], m.bx[x]);
#pragma unroll
    for (int j = 0; j < FH; j++) {
        const int zl = zg + j, z = z0 + zl;
        if (!inbox(d, x, z)) continue;
        const long long i = (long long)z * d.X + x;
        float v[8], w[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            v[q] = sx[zl][tx + q];
            w[q] = sz[zl + q][tx];
        }
        const float v0 = dsum(v) / m.dx;
        const float mxt = kx ? m.bx[x] * mxxp[i] + m.ax[x] * v0 : 0.0f;
        if (kx) mxxc[i] = mxt;
        const float vxx = v0 / m.Kx[x] + mxt;
        const float w0 = dsum(w) / m.dz;
        const bool kz = live(m.az[z], m.bz[z]);
        const float mzt = kz ? m.bz[z] * mzzp[i] + m.az[z] * w0 : 0.0f;
        if (kz) mzzc[i] = mzt;
        const float vzz = w0 / m.Kz[z] + mzt;
        const float lap = vxx + vzz;
        const float pm2 = P[((t - 2) % 4) * pl + i];
        P[(t % 4) * pl + i] = 2.0f * pm1[j] - pm2 + lap * m.dt2 * m.kappa[i];
    }
}

// Sources of shot s (in source order, on one lane) at t >= 2, then every
// receiver: rec[s][r][t] = P(t)[rx r].  shot_idx[s]: the active source when
// ppe == 1 (-1: all); every source when ppe != 1.
__global__ void cpml_src_rec(Fields f, Medium m, Dims d, int t, int T, int ppe, int n_src,
                             const int *sx, const int *sz, const float *src, const int *shot_idx,
                             int n_rx, const int *rx, const int *rz, float *rec)
{
    const int s = blockIdx.x;
    float *Pt = f.P + ((long long)s * 4 + t % 4) * d.plane;
    if (t > 1 && threadIdx.x == 0) {
        const int idx = shot_idx[s];
        for (int n = 0; n < n_src; n++) {
            if (ppe == 1 && idx != -1 && idx != n) continue;
            const long long i = (long long)sz[n] * d.X + sx[n];
            Pt[i] += m.qt_pi * m.cq[i] * src[(long long)n * T + t] * m.dt2;
        }
    }
    __syncthreads();
    for (int r = threadIdx.x; r < n_rx; r += blockDim.x)
        rec[((long long)s * n_rx + r) * T + t] = Pt[(long long)rz[r] * d.X + rx[r]];
}

template <typename T>
int dalloc(T **p, size_t n)
{
    *p = nullptr;
    if (n == 0) return 0;
    if (hipMalloc((void **)p, n * sizeof(T)) != hipSuccess) {
        *p = nullptr;
        return cfail(-3, "device allocation of %zu bytes failed", n * sizeof(T));
    }
    return 0;
}
template <typename T>
void dfree(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

}  // namespace

struct otfwi_cpml {
    int X = 0, Z = 0, T = 0, ppe = 1;
    float dx = 1, dz = 1, dt2 = 1, qt_pi = 1;
    float *med = nullptr;        // cq, kappa, rho_hx, rho_hz: 4 planes
    float *prof = nullptr;       // Kx ax bx Kxh axh bxh [X] each, then the same [Z]
    float *init = nullptr;       // [2][plane]
    int n_src = 0, n_rx = 0;
    int *pts = nullptr;          // sx, sz, rx, rz
    float *src = nullptr;        // [n_src][T]
    int ntx = 0, ntz = 0;            // tiles along x and z
    int cap = 0;                 // shots the state buffers hold
    float *state = nullptr;      // P 4 + mx, mz, mxx, mzz 2 each: 12 planes per shot
    float *rec = nullptr;        // [cap][n_rx][T]
    int *shot_idx = nullptr;     // [cap]
    hipStream_t stream = nullptr;
};

namespace {

size_t plane_of(const otfwi_cpml *h) { return (size_t)h->X * h->Z; }

Medium medium(const otfwi_cpml *h)
{
    const size_t pl = plane_of(h);
    Medium m;
    m.cq = h->med; m.kappa = h->med + pl; m.rho_hx = h->med + 2 * pl; m.rho_hz = h->med + 3 * pl;
    const float *px = h->prof, *pz = h->prof + 6 * (size_t)h->X;
    m.Kx = px; m.ax = px + h->X; m.bx = px + 2 * h->X;
    m.Kxh = px + 3 * h->X; m.axh = px + 4 * h->X; m.bxh = px + 5 * h->X;
    m.Kz = pz; m.az = pz + h->Z; m.bz = pz + 2 * h->Z;
    m.Kzh = pz + 3 * h->Z; m.azh = pz + 4 * h->Z; m.bzh = pz + 5 * h->Z;
    m.dx = h->dx; m.dz = h->dz; m.dt2 = h->dt2; m.qt_pi = h->qt_pi;
    return m;
}

int ensure_state(otfwi_cpml *h, int n)
{
    if (n <= h->cap && h->state) return 0;
    dfree(h->state); dfree(h->rec); dfree(h->shot_idx);
    h->cap = 0;
    if (int rc = dalloc(&h->state, (size_t)n * 12 * plane_of(h))) return rc;
    if (int rc = dalloc(&h->rec, (size_t)n * std::max(h->n_rx, 1) * h->T)) return rc;
    if (int rc = dalloc(&h->shot_idx, (size_t)n)) return rc;
    h->cap = n;
    return 0;
}

}  // namespace

extern "C" {

const char *otfwi_cpml_last_error(void) { return g_err.c_str(); }

int otfwi_cpml_create(otfwi_cpml **out, int X, int Z, int T, int ppe, const float *consts,
                      const float *const *fields, const float *const *prof, const float *initial)
{
    if (!out || X < 9 || Z < 9 || T < 1 || !consts || !fields || !prof)
        return cfail(-1, "cpml_create: bad arguments (X=%d Z=%d T=%d)", X, Z, T);
    for (int k = 0; k < 4; k++)
        if (!fields[k]) return cfail(-1, "cpml_create: medium field %d is NULL", k);
    for (int k = 0; k < 12; k++)
        if (!prof[k]) return cfail(-1, "cpml_create: profile %d is NULL", k);
    auto *h = new otfwi_cpml;
    h->X = X; h->Z = Z; h->T = T; h->ppe = ppe;
    h->dx = consts[0]; h->dz = consts[1]; h->dt2 = consts[3]; h->qt_pi = consts[4];
    const size_t pl = plane_of(h);
    int rc = 0;
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) rc = cfail(-2, "stream: %s", hipGetErrorString(e));
    if (!rc) rc = dalloc(&h->med, 4 * pl);
    if (!rc) rc = dalloc(&h->prof, (size_t)6 * (X + Z));
    if (!rc) rc = dalloc(&h->init, 2 * pl);
    for (int k = 0; !rc && k < 4; k++)
        if (hipMemcpy(h->med + k * pl, fields[k], pl * sizeof(float), hipMemcpyHostToDevice))
            rc = cfail(-2, "medium upload failed");
    for (int k = 0; !rc && k < 12; k++) {
        const int n = k < 6 ? X : Z;
        float *dst = h->prof + (k < 6 ? (size_t)k * X : (size_t)6 * X + (size_t)(k - 6) * Z);
        if (hipMemcpy(dst, prof[k], n * sizeof(float), hipMemcpyHostToDevice))
            rc = cfail(-2, "profile upload failed");
    }
    h->ntx = (X + FX - 1) / FX;
    h->ntz = (Z + FZ - 1) / FZ;
    if (!rc) {
        if (initial)
            e = hipMemcpy(h->init, initial, 2 * pl * sizeof(float), hipMemcpyHostToDevice);
        else
            e = hipMemset(h->init, 0, 2 * pl * sizeof(float));
        if (e != hipSuccess) rc = cfail(-2, "initial state upload failed");
    }
    if (rc) {
        otfwi_cpml_destroy(h);
        return rc;
    }
    *out = h;
    return 0;
}

int otfwi_cpml_destroy(otfwi_cpml *h)
{
    if (!h) return 0;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    dfree(h->med); dfree(h->prof); dfree(h->init); dfree(h->pts); dfree(h->src);
    dfree(h->state); dfree(h->rec); dfree(h->shot_idx);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return 0;
}

int otfwi_cpml_set_cquad(otfwi_cpml *h, const float *cq)
{
    if (!h || !cq) return cfail(-1, "set_cquad: bad arguments");
    CHK(hipMemcpy(h->med, cq, plane_of(h) * sizeof(float), hipMemcpyHostToDevice));
    return 0;
}

int otfwi_cpml_set_points(otfwi_cpml *h, int n_src, const int *sx, const int *sz,
                          const float *src, int n_rx, const int *rx, const int *rz)
{
    if (!h || n_src < 0 || n_rx < 0 || (n_src && (!sx || !sz || !src)) || (n_rx && (!rx || !rz)))
        return cfail(-1, "set_points: bad arguments");
    for (int k = 0; k < n_src; k++)
        if (sx[k] < 0 || sx[k] >= h->X || sz[k] < 0 || sz[k] >= h->Z)
            return cfail(-1, "source %d at (%d, %d) is outside the grid", k, sx[k], sz[k]);
    for (int k = 0; k < n_rx; k++)
        if (rx[k] < 0 || rx[k] >= h->X || rz[k] < 0 || rz[k] >= h->Z)
            return cfail(-1, "receiver %d at (%d, %d) is outside the grid", k, rx[k], rz[k]);
    dfree(h->pts); dfree(h->src);
    if (h->n_rx != n_rx) { dfree(h->state); dfree(h->rec); dfree(h->shot_idx); h->cap = 0; }
    h->n_src = n_src; h->n_rx = n_rx;
    const int np = 2 * (n_src + n_rx);
    if (int rc = dalloc(&h->pts, (size_t)std::max(np, 1))) return rc;
    std::vector<int> pts;
    pts.insert(pts.end(), sx, sx + n_src);
    pts.insert(pts.end(), sz, sz + n_src);
    pts.insert(pts.end(), rx, rx + n_rx);
    pts.insert(pts.end(), rz, rz + n_rx);
    if (np) CHK(hipMemcpy(h->pts, pts.data(), np * sizeof(int), hipMemcpyHostToDevice));
    if (n_src) {
        if (int rc = dalloc(&h->src, (size_t)n_src * h->T)) return rc;
        CHK(hipMemcpy(h->src, src, (size_t)n_src * h->T * sizeof(float), hipMemcpyHostToDevice));
    }
    return 0;
}

int otfwi_cpml_forward(otfwi_cpml *h, int n_shots, const int *shot_idx, float *rec_out)
{
    if (!h || n_shots <= 0 || !shot_idx) return cfail(-1, "cpml_forward: bad arguments");
    if (!h->pts) return cfail(-4, "cpml_forward: sources and receivers not set");
    for (int s = 0; s < n_shots; s++)   // an index past the last source fires none, as in the reference
        if (shot_idx[s] < -1) return cfail(-1, "shot %d: source index %d", s, shot_idx[s]);
    if (int rc = ensure_state(h, n_shots)) return rc;
    const size_t pl = plane_of(h);
    hipStream_t st = h->stream;
    CHK(hipMemcpyAsync(h->shot_idx, shot_idx, n_shots * sizeof(int), hipMemcpyHostToDevice, st));
    // P slots 0, 1 = initial, 2, 3 = 0; memory variables 0
    CHK(hipMemsetAsync(h->state, 0, (size_t)n_shots * 12 * pl * sizeof(float), st));
    Fields f;
    f.P = h->state;
    float *rest = h->state + (size_t)n_shots * 4 * pl;
    f.mx = rest; f.mz = rest + (size_t)n_shots * 2 * pl;
    f.mxx = rest + (size_t)n_shots * 4 * pl; f.mzz = rest + (size_t)n_shots * 6 * pl;
    for (int s = 0; s < n_shots; s++)
        CHK(hipMemcpyAsync(f.P + (size_t)s * 4 * pl, h->init, 2 * pl * sizeof(float),
                           hipMemcpyDeviceToDevice, st));
    const Medium m = medium(h);
    const Dims d{h->X, h->Z, (long long)pl};
    const dim3 grid(h->ntx, h->ntz, n_shots), blk(256);
    const int *sx = h->pts, *sz = h->pts + h->n_src, *rx = h->pts + 2 * h->n_src;
    const int *rz = rx + h->n_rx;
    for (int t = 0; t < h->T; t++) {
        if (t > 1) hipLaunchKernelGGL(cpml_step, grid, blk, 0, st, f, m, d, t);
        hipLaunchKernelGGL(cpml_src_rec, dim3(n_shots), dim3(64), 0, st, f, m, d, t, h->T, h->ppe,
                           h->n_src, sx, sz, h->src, h->shot_idx, h->n_rx, rx, rz, h->rec);
    }
    CHK(hipGetLastError());
    if (rec_out && h->n_rx)
        CHK(hipMemcpyAsync(rec_out, h->rec, (size_t)n_shots * h->n_rx * h->T * sizeof(float),
                           hipMemcpyDeviceToHost, st));
    CHK(hipStreamSynchronize(st));
    return 0;
}

// ---------------------------------------------------------------------------
// The reference's C ABI (regression_CPML.cu:349-501): one process-wide model,
// void functions, the recording in a library-owned pinned buffer.
// ---------------------------------------------------------------------------
static otfwi_cpml *g_model = nullptr;
static float *g_rec_h = nullptr;
typedef double take3double1int_givedouble(float *, float *, float *, float *, int, int);
static take3double1int_givedouble *g_adj = nullptr;

void def_python_adj(take3double1int_givedouble *func) { g_adj = func; }

void init_memory_sim(int x, int z, int t, int ns, int nm, int ppe, int *ps_x, int *ps_z,
                     int *pm_x, int *pm_z, float dx_h, float dz_h, float dt_h, float dt2_h,
                     float qtpi_h, float *cq, float *kp_unr, float *rho_h, float *rho_hx_h,
                     float *rho_hz_h, float *d_x_h, float *K_x_h, float *alpha_x_h,
                     float *a_x_h, float *b_x_h, float *d_x_half_h, float *K_x_half_h,
                     float *alpha_x_half_h, float *a_x_half_h, float *b_x_half_h, float *d_z_h,
                     float *K_z_h, float *alpha_z_h, float *a_z_h, float *b_z_h,
                     float *d_z_half_h, float *K_z_half_h, float *alpha_z_half_h,
                     float *a_z_half_h, float *b_z_half_h, float *src, float *init, float **rec)
{
    (void)rho_h; (void)d_x_h; (void)alpha_x_h; (void)d_x_half_h; (void)alpha_x_half_h;
    (void)d_z_h; (void)alpha_z_h; (void)d_z_half_h; (void)alpha_z_half_h;
    otfwi_cpml_destroy(g_model);
    g_model = nullptr;
    if (g_rec_h) (void)hipHostFree(g_rec_h);
    g_rec_h = nullptr;
    const float consts[5] = {dx_h, dz_h, dt_h, dt2_h, qtpi_h};
    const float *fields[4] = {cq, kp_unr, rho_hx_h, rho_hz_h};
    const float *prof[12] = {K_x_h, a_x_h, b_x_h, K_x_half_h, a_x_half_h, b_x_half_h,
                             K_z_h, a_z_h, b_z_h, K_z_half_h, a_z_half_h, b_z_half_h};
    if (otfwi_cpml_create(&g_model, x, z, t, ppe, consts, fields, prof, init) ||
        otfwi_cpml_set_points(g_model, ns, ps_x, ps_z, src, nm, pm_x, pm_z)) {
        fprintf(stderr, "init_memory_sim: %s\n", g_err.c_str());
        return;
    }
    if (hipHostMalloc((void **)&g_rec_h, (size_t)std::max(nm, 1) * t * sizeof(float), 0) !=
        hipSuccess) {
        fprintf(stderr, "init_memory_sim: pinned recording buffer allocation failed\n");
        g_rec_h = nullptr;
    }
    if (rec) *rec = g_rec_h;
}

void setCquad(float *cq)
{
    if (g_model && otfwi_cpml_set_cquad(g_model, cq))
        fprintf(stderr, "setCquad: %s\n", g_err.c_str());
}

void set_source(int ns, int *ps_x, int *ps_z)
{
    // new source positions, same wavelets and receivers (regression_CPML.cu:429-436);
    // the first ns wavelets stay in use
    if (!g_model) return;
    std::vector<int> rx(g_model->n_rx), rz(g_model->n_rx);
    std::vector<float> src((size_t)std::max(ns, 0) * g_model->T);
    if (g_model->n_rx) {
        std::vector<int> pts(2 * (g_model->n_src + g_model->n_rx));
        if (hipMemcpy(pts.data(), g_model->pts, pts.size() * sizeof(int), hipMemcpyDeviceToHost))
            return;
        std::copy(pts.begin() + 2 * g_model->n_src, pts.begin() + 2 * g_model->n_src + g_model->n_rx,
                  rx.begin());
        std::copy(pts.begin() + 2 * g_model->n_src + g_model->n_rx, pts.end(), rz.begin());
    }
    const int keep = std::min(ns, g_model->n_src);
    if (keep > 0 && hipMemcpy(src.data(), g_model->src, (size_t)keep * g_model->T * sizeof(float),
                              hipMemcpyDeviceToHost))
        return;
    if (otfwi_cpml_set_points(g_model, ns, ps_x, ps_z, src.data(), g_model->n_rx, rx.data(),
                              rz.data()))
        fprintf(stderr, "set_source: %s\n", g_err.c_str());
}

void cuda_simulate2(int en_out, int idx_source)
{
    (void)en_out;   // video frames are not produced
    if (!g_model || !g_rec_h) return;
    if (otfwi_cpml_forward(g_model, 1, &idx_source, g_rec_h))
        fprintf(stderr, "cuda_simulate2: %s\n", g_err.c_str());
}

}  // extern "C"
