// Exact discrete adjoint of the forward recursion of prop.hip.
//
// Forward (one time step t, interior cells 4 <= x < X-4, 4 <= z < Z-4, fields
// read as 0 elsewhere; m = c^2 dt/dx, cA = dt/dx):
//   Px_t = kx Px_{t-1} + m D-x Ax_{t-1},   Pz_t = kz Pz_{t-1} + m D-z Az_{t-1}
//   (sources add constants; receivers record r_t = Px_t + Pz_t)
//   Ax_t = kx Ax_{t-1} + cA D+x (Px_t + Pz_t),   Az_t likewise with kz, D+z.
// With the masked operators, (D+)^T = -D- and (D-)^T = -D+, so the adjoint
// variables u, w (of Px, Pz) and a, b (of Ax, Az) run backwards, t = T-1..1:
//   u_t = kx u_{t+1} - cA (D-x a_t + D-z b_t) + rho_t      (rho = dJ/dr at receivers)
//   w_t = kz w_{t+1} - cA (D-x a_t + D-z b_t) + rho_t
//   dJ/dm += u_t D-x Ax_{t-1} + w_t D-z Az_{t-1}
//   a_{t-1} = kx a_t - D+x (m u_t),   b_{t-1} = kz b_t - D+z (m w_t)
// from u_T = w_T = 0, a_{T-1} = b_{T-1} = 0.  The forward A fields come from
// snapshot checkpointing (api.hip).  Fields use the padded layout of prop.hip
// (zero padding and zero outside the updated box), so stencils need no masks.
#include "otfwi_internal.h"

namespace otfwi {

namespace {

constexpr double C1 = 1.196289e+00, C2 = -7.975260e-02, C3 = 9.570313e-03, C4 = -6.975446e-04;

// D- f at i: sum_k c_k (f[i + (k-1) s] - f[i - k s])
template <typename R>
__device__ __forceinline__ R dminus(const R *f, long long i, long long s)
{
    R v = (R)C4 * (f[i + 3 * s] - f[i - 4 * s]);
    v = fma((R)C3, f[i + 2 * s] - f[i - 3 * s], v);
    v = fma((R)C2, f[i + s] - f[i - 2 * s], v);
    v = fma((R)C1, f[i] - f[i - s], v);
    return v;
}

// D+ (m f) at i: sum_k c_k (m f)[i + k s] - (m f)[i + (1-k) s]; m indexed by
// the shot-independent plane offset p, f by the field offset i.
template <typename R>
__device__ __forceinline__ R dplus_mf(const R *m, long long p, const R *f, long long i,
                                      long long s)
{
    auto g = [&](long long o) { return m[p + o] * f[i + o]; };
    R v = (R)C4 * (g(4 * s) - g(-3 * s));
    v = fma((R)C3, g(3 * s) - g(-2 * s), v);
    v = fma((R)C2, g(2 * s) - g(-s), v);
    v = fma((R)C1, g(s) - g(0), v);
    return v;
}

template <typename R>
__global__ void __launch_bounds__(256) xadj_u_kernel(XAdjArgs<R> a)
{
    const int x = BOX + blockIdx.x * 64 + threadIdx.x;
    const int z = BOX + blockIdx.y * 4 + threadIdx.y;
    const int s = blockIdx.z;
    if (x >= a.X - BOX || z >= a.Z - BOX) return;
    const long long ld = a.ld;
    const long long p = (long long)(z + ZOFF) * ld + x + XOFF;
    const long long i = (long long)s * a.sstride + p;
    const R dA = dminus(a.a, i, 1) + dminus(a.b, i, ld);
    R r = R(0);
    const int ri = a.rxmap[p];
    if (ri >= 0) r = a.rho[((long long)s * a.n_rx + ri) * a.T + a.t];
    const R kx = a.kx[XOFF + x], kz = a.kz[ZOFF + z];
    const R u = kx * a.u0[i] - a.cA * dA + r;
    const R w = kz * a.w0[i] - a.cA * dA + r;
    a.u1[i] = u;
    a.w1[i] = w;
    a.grad[i] += u * dminus(a.Ax, i, 1) + w * dminus(a.Az, i, ld);
}

template <typename R>
__global__ void __launch_bounds__(256) xadj_a_kernel(XAdjArgs<R> a)
{
    const int x = BOX + blockIdx.x * 64 + threadIdx.x;
    const int z = BOX + blockIdx.y * 4 + threadIdx.y;
    const int s = blockIdx.z;
    if (x >= a.X - BOX || z >= a.Z - BOX) return;
    const long long ld = a.ld;
    const long long p = (long long)(z + ZOFF) * ld + x + XOFF;
    const long long i = (long long)s * a.sstride + p;
    const R kx = a.kx[XOFF + x], kz = a.kz[ZOFF + z];
    a.a1[i] = kx * a.a[i] - dplus_mf(a.cm, p, a.u1, i, 1);
    a.b1[i] = kz * a.b[i] - dplus_mf(a.cm, p, a.w1, i, ld);
}

}  // namespace

template <typename R>
hipError_t launch_xadj_step(const XAdjArgs<R> &a, hipStream_t st)
{
    const int nx = a.X - 2 * BOX, nz = a.Z - 2 * BOX;
    if (nx <= 0 || nz <= 0 || a.nshots <= 0) return hipSuccess;
    dim3 grid((nx + 63) / 64, (nz + 3) / 4, a.nshots), block(64, 4);
    hipLaunchKernelGGL((xadj_u_kernel<R>), grid, block, 0, st, a);
    hipLaunchKernelGGL((xadj_a_kernel<R>), grid, block, 0, st, a);
    return hipGetLastError();
}
template hipError_t launch_xadj_step<float>(const XAdjArgs<float> &, hipStream_t);
template hipError_t launch_xadj_step<double>(const XAdjArgs<double> &, hipStream_t);

}  // namespace otfwi
