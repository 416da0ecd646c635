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
// Work-groups of 64 x 16 cells stage their stencil inputs in LDS.
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
    float *vx1, *vz1;  // [S][plane]
};

struct Medium {
    const float *cq, *kappa, *rho_hx, *rho_hz;           // [plane]
    const float *Kx, *ax, *bx, *Kxh, *axh, *bxh;         // [X]
    const float *Kz, *az, *bz, *Kzh, *azh, *bzh;         // [Z]
    float dx, dz, dt2, qt_pi;
};

// A work-group of 256 threads owns a TX x TZ tile (64 x 16): it stages the
// stencil inputs of the tile (P(t-1) with a 4-cell halo for frame 1; the x
// derivative field with an x halo and the z one with a z halo for frame 2)
// in LDS once, with the box rule applied while loading; each thread then
// computes 4 cells of one column.  skip*: tiles whose profile values are all
// a = 0, b = 1 (the interior): there the memory variables start at 0 and
// stay exactly 0, so they are neither read nor written.
constexpr int TX = 64, TZ = 16, TH = 4;

// derivPMLdx / derivPMLdz (deriv_macros.h:85-108) of 8 staged values, the
// reference's term order: v[0..8) = taps -3-off .. 4-off
__device__ __forceinline__ float dsum(const float *v)
{
    return (float)(-C4 * (double)v[7] + C3 * (double)v[6] - C2 * (double)v[5] +
                   C1 * (double)v[4] - C1 * (double)v[3] + C2 * (double)v[2] -
                   C3 * (double)v[1] + C4 * (double)v[0]);
}

__global__ void __launch_bounds__(256)
cpml_frame1(Fields f, Medium m, Dims d, int t, const unsigned char *skipx,
            const unsigned char *skipz)
{
    __shared__ float sp[TZ + 8][TX + 8];
    const int tid = threadIdx.x, s = blockIdx.z;
    const int x0 = blockIdx.x * TX, z0 = blockIdx.y * TZ;
    const long long pl = d.plane;
    const float *P1 = f.P + ((long long)s * 4 + (t - 1) % 4) * pl;
    for (int k = tid; k < (TZ + 8) * (TX + 8); k += 256) {
        const int r = k / (TX + 8), c = k - r * (TX + 8);
        sp[r][c] = boxv(P1, d, x0 - 4 + c, z0 - 4 + r);
    }
    __syncthreads();
    const int tx = tid & (TX - 1), zg = (tid / TX) * TH;
    const int x = x0 + tx;
    const bool kx = !skipx[blockIdx.x], kz = !skipz[blockIdx.y];
    const int c = t & 1, p = c ^ 1;
    float *mx = f.mx + (long long)s * 2 * pl, *mz = f.mz + (long long)s * 2 * pl;
#pragma unroll
    for (int j = 0; j < TH; j++) {
        const int zl = zg + j, z = z0 + zl;
        if (!inbox(d, x, z)) continue;
        const long long i = (long long)z * d.X + x;
        float v[8], w[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {   // taps -3 .. 4 (offset 0)
            v[k] = sp[zl + 4][tx + 4 - 3 + k];
            w[k] = sp[zl + 4 - 3 + k][tx + 4];
        }
        const float v0 = dsum(v) / m.dx;
        const float mxt = kx ? m.bxh[x] * mx[p * pl + i] + m.axh[x] * v0 : 0.0f;
        if (kx) mx[c * pl + i] = mxt;
        f.vx1[s * pl + i] = (v0 / m.Kxh[x] + mxt) / m.rho_hx[i];
        const float w0 = dsum(w) / m.dz;
        const float mzt = kz ? m.bzh[z] * mz[p * pl + i] + m.azh[z] * w0 : 0.0f;
        if (kz) mz[c * pl + i] = mzt;
        f.vz1[s * pl + i] = (w0 / m.Kzh[z] + mzt) / m.rho_hz[i];
    }
}

__global__ void __launch_bounds__(256)
cpml_frame2(Fields f, Medium m, Dims d, int t, const unsigned char *skipx,
            const unsigned char *skipz)
{
    __shared__ float sx[TZ][TX + 8];
    __shared__ float sz[TZ + 8][TX];
    const int tid = threadIdx.x, s = blockIdx.z;
    const int x0 = blockIdx.x * TX, z0 = blockIdx.y * TZ;
    const long long pl = d.plane;
    const float *VX = f.vx1 + s * pl, *VZ = f.vz1 + s * pl;
    for (int k = tid; k < TZ * (TX + 8); k += 256) {
        const int r = k / (TX + 8), c = k - r * (TX + 8);
        sx[r][c] = boxv(VX, d, x0 - 4 + c, z0 + r);
    }
    for (int k = tid; k < (TZ + 8) * TX; k += 256) {
        const int r = k / TX, c = k - r * TX;
        sz[r][c] = boxv(VZ, d, x0 + c, z0 - 4 + r);
    }
    __syncthreads();
    const int tx = tid & (TX - 1), zg = (tid / TX) * TH;
    const int x = x0 + tx;
    const bool kx = !skipx[blockIdx.x], kz = !skipz[blockIdx.y];
    const int c = t & 1, p = c ^ 1;
    float *mxx = f.mxx + (long long)s * 2 * pl, *mzz = f.mzz + (long long)s * 2 * pl;
    float *P = f.P + (long long)s * 4 * pl;
#pragma unroll
    for (int j = 0; j < TH; j++) {
        const int zl = zg + j, z = z0 + zl;
        if (!inbox(d, x, z)) continue;
        const long long i = (long long)z * d.X + x;
        float v[8], w[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {   // taps -4 .. 3 (offset 1)
            v[k] = sx[zl][tx + 4 - 4 + k];
            w[k] = sz[zl + 4 - 4 + k][tx];
        }
        const float v0 = dsum(v) / m.dx;
        const float mxt = kx ? m.bx[x] * mxx[p * pl + i] + m.ax[x] * v0 : 0.0f;
        if (kx) mxx[c * pl + i] = mxt;
        const float vxx = v0 / m.Kx[x] + mxt;
        const float w0 = dsum(w) / m.dz;
        const float mzt = kz ? m.bz[z] * mzz[p * pl + i] + m.az[z] * w0 : 0.0f;
        if (kz) mzz[c * pl + i] = mzt;
        const float vzz = w0 / m.Kz[z] + mzt;
        const float lap = vxx + vzz;
        const float p1 = P[((t - 1) % 4) * pl + i], p2 = P[((t - 2) % 4) * pl + i];
        P[(t % 4) * pl + i] = 2.0f * p1 - p2 + lap * m.dt2 * m.kappa[i];
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
    unsigned char *skip = nullptr;   // per tile column / row: frame 1 x, z; frame 2 x, z
    int ntx = 0, ntz = 0;
    int cap = 0;                 // shots the state buffers hold
    float *state = nullptr;      // P 4 + mx, mz, mxx, mzz 2 each + vx1, vz1: 14 planes per shot
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
    if (int rc = dalloc(&h->state, (size_t)n * 14 * plane_of(h))) return rc;
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
    if (!rc) {
        // tiles whose memory variables stay 0: a == 0 and b == 1 on every
        // column (row) of the tile, for the half-point (frame 1) and the
        // full-point (frame 2) profiles
        h->ntx = (X + TX - 1) / TX;
        h->ntz = (Z + TZ - 1) / TZ;
        std::vector<unsigned char> sk(2 * (h->ntx + h->ntz), 1);
        auto still = [](const float *a, const float *b, int lo, int hi) {
            for (int k = lo; k < hi; k++)
                if (a[k] != 0.0f || b[k] != 1.0f) return false;
            return true;
        };
        for (int b = 0; b < h->ntx; b++) {
            const int lo = b * TX, hi = std::min(X, lo + TX);
            sk[b] = still(prof[4], prof[5], lo, hi);                          // axh, bxh
            sk[h->ntx + h->ntz + b] = still(prof[1], prof[2], lo, hi);        // ax, bx
        }
        for (int b = 0; b < h->ntz; b++) {
            const int lo = b * TZ, hi = std::min(Z, lo + TZ);
            sk[h->ntx + b] = still(prof[10], prof[11], lo, hi);               // azh, bzh
            sk[2 * h->ntx + h->ntz + b] = still(prof[7], prof[8], lo, hi);    // az, bz
        }
        rc = dalloc(&h->skip, sk.size());
        if (!rc && hipMemcpy(h->skip, sk.data(), sk.size(), hipMemcpyHostToDevice))
            rc = cfail(-2, "tile flags upload failed");
    }
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
    dfree(h->med); dfree(h->prof); dfree(h->init); dfree(h->pts); dfree(h->src); dfree(h->skip);
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
    // P slots 0, 1 = initial, 2, 3 = 0; memory variables and derivatives 0
    CHK(hipMemsetAsync(h->state, 0, (size_t)n_shots * 14 * pl * sizeof(float), st));
    Fields f;
    f.P = h->state;
    float *rest = h->state + (size_t)n_shots * 4 * pl;
    f.mx = rest; f.mz = rest + (size_t)n_shots * 2 * pl;
    f.mxx = rest + (size_t)n_shots * 4 * pl; f.mzz = rest + (size_t)n_shots * 6 * pl;
    f.vx1 = rest + (size_t)n_shots * 8 * pl; f.vz1 = rest + (size_t)n_shots * 9 * pl;
    for (int s = 0; s < n_shots; s++)
        CHK(hipMemcpyAsync(f.P + (size_t)s * 4 * pl, h->init, 2 * pl * sizeof(float),
                           hipMemcpyDeviceToDevice, st));
    const Medium m = medium(h);
    const Dims d{h->X, h->Z, (long long)pl};
    const dim3 grid(h->ntx, h->ntz, n_shots), blk(256);
    const unsigned char *sk1x = h->skip, *sk1z = h->skip + h->ntx;
    const unsigned char *sk2x = h->skip + h->ntx + h->ntz, *sk2z = sk2x + h->ntx;
    const int *sx = h->pts, *sz = h->pts + h->n_src, *rx = h->pts + 2 * h->n_src;
    const int *rz = rx + h->n_rx;
    for (int t = 0; t < h->T; t++) {
        if (t > 1) {
            hipLaunchKernelGGL(cpml_frame1, grid, blk, 0, st, f, m, d, t, sk1x, sk1z);
            hipLaunchKernelGGL(cpml_frame2, grid, blk, 0, st, f, m, d, t, sk2x, sk2z);
        }
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
