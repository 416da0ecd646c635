// Misfit kernels: per-trace L2 and 1D Wasserstein-2 misfits with their
// adjoint sources, the running sum of the time-flipped adjoint source that
// the adjoint step injects, and the FP64 cross-shot reductions.
//
// W2 follows adjoint_functions.wasserstein (adjoint_functions.py:25-67) and
// the scipy semantics it relies on:
//   t_k = k/(T-1) (numpy linspace, last sample exactly 1), dt = t_1 - t_0;
//   c = max(-1.1 min(obs), 0.1 max(obs)); g = obs + c, f = sim + c;
//   g /= sum(g dt), f /= sum(f dt);
//   G = [0, cumsum(diff(t) (g[1:] + g[:-1]) / 2)], F likewise;
//   interp1d(xs, ys, 'linear', extrapolate): stable sort of xs,
//     i = clip(searchsorted_left(xs, q), 1, n-1),
//     y = (ys[i]-ys[i-1]) / (xs[i]-xs[i-1]) * (q - xs[i-1]) + ys[i-1];
//   q = G^-1(F(t)); D = 2 [0, cumsum((h[1:] + h[:-1]) / 2)] with h = t - q
//     (unit spacing); adj = D(t) - D(1); W2 = sum (F^-1(t) - G^-1(t))^2.
// One wavefront owns one trace; CDFs are built with FP64 wavefront prefix
// scans and kept in LDS; inverse CDFs are binary searches in LDS.  A CDF that
// is not monotone (sim + c < 0 somewhere) is stably sorted into a global
// scratch area first (rank sort: O(T^2) but only for such traces).
#include "otfwi_internal.h"

#include <cmath>

namespace otfwi {

constexpr int WAVE = 64;

__device__ __forceinline__ double tval(int k, int T, double step)
{
    return k == T - 1 ? 1.0 : (double)k * step;
}

__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = WAVE / 2; o > 0; o >>= 1)
        v += __shfl_xor(v, o, WAVE);
    return v;
}
__device__ __forceinline__ double wave_min(double v)
{
    for (int o = WAVE / 2; o > 0; o >>= 1)
        v = fmin(v, __shfl_xor(v, o, WAVE));
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
    for (int o = WAVE / 2; o > 0; o >>= 1)
        v = fmax(v, __shfl_xor(v, o, WAVE));
    return v;
}
// exclusive prefix sum over the wavefront (lane order)
__device__ __forceinline__ double wave_excl_scan(double v)
{
    const int lane = threadIdx.x & (WAVE - 1);
    double incl = v;
    for (int o = 1; o < WAVE; o <<= 1) {
        const double u = __shfl_up(incl, o, WAVE);
        if (lane >= o) incl += u;
    }
    return incl - v;
}

// Trapezoid running integral, in place: y[k] <- [0, cumsum(d_k (y[k+1]+y[k])/2)]
// with d_k = t_{k+1} - t_k (spacing from tval) or 1 (unit).  Lane l owns the
// contiguous chunk [l*C, (l+1)*C); the only cross-chunk operand, y[l*C-1], is
// read before any lane writes.
template <bool UNIT>
__device__ void cumtrapz_inplace(double *y, int T, double step)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const int C = (T + WAVE - 1) / WAVE;
    const int k0 = lane * C, k1 = min(k0 + C, T);
    const double left = (k0 > 0 && k0 < T) ? y[k0 - 1] : 0.0;
    double local = 0.0;
    for (int k = max(k0, 1); k < k1; k++) {
        const double d = UNIT ? 1.0 : tval(k, T, step) - tval(k - 1, T, step);
        const double ym = (k == k0) ? left : y[k - 1];
        local += d * (y[k] + ym) / 2.0;
    }
    const double off = wave_excl_scan(local);
    __syncthreads();
    double run = off, prev = left;
    for (int k = k0; k < k1; k++) {
        const double yk = y[k];
        if (k > 0) {
            const double d = UNIT ? 1.0 : tval(k, T, step) - tval(k - 1, T, step);
            run += d * (yk + prev) / 2.0;
        }
        prev = yk;
        y[k] = (k == 0) ? 0.0 : run;
    }
    __syncthreads();
}

// searchsorted(xs, q, side='left') on a sorted array of n values.
template <typename XS>
__device__ __forceinline__ int lower_bound(const XS &xs, int n, double q)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (xs(mid) < q) lo = mid + 1; else hi = mid;
    }
    return lo;
}

template <typename XS, typename YS>
__device__ __forceinline__ double interp_lin(const XS &xs, const YS &ys, int n, double q)
{
    int i = lower_bound(xs, n, q);
    i = i < 1 ? 1 : (i > n - 1 ? n - 1 : i);
    const double xl = xs(i - 1), xh = xs(i), yl = ys(i - 1), yh = ys(i);
    const double slope = (yh - yl) / (xh - xl);
    return slope * (q - xl) + yl;
}

// Value of interp1d(t, Y) at t_k itself (searchsorted gives k).
__device__ __forceinline__ double interp_at_grid(const double *Y, int k, int T, double step)
{
    const int i = k < 1 ? 1 : k;
    const double xl = tval(i - 1, T, step), xh = tval(i, T, step);
    const double slope = (Y[i] - Y[i - 1]) / (xh - xl);
    return slope * (tval(k, T, step) - xl) + Y[i - 1];
}

// Stable sort of (xs, t) by xs into (sx, st): element j goes to position
// #{m : xs_m < xs_j} + #{m < j : xs_m == xs_j}, the order numpy's mergesort
// argsort produces.
__device__ void rank_sort(const double *xs, int T, double step, double *sx, double *st)
{
    const int lane = threadIdx.x & (WAVE - 1);
    for (int j = lane; j < T; j += WAVE) {
        const double v = xs[j];
        int r = 0;
        for (int m = 0; m < T; m++) {
            const double u = xs[m];
            r += (u < v) || (u == v && m < j);
        }
        sx[r] = v;
        st[r] = tval(j, T, step);
    }
    __syncthreads();
}

__device__ bool is_monotone(const double *Y, int T)
{
    const int lane = threadIdx.x & (WAVE - 1);
    int bad = 0;
    for (int k = lane + 1; k < T; k += WAVE)
        bad |= Y[k] < Y[k - 1];
    return __any(bad) == 0;
}

template <typename R, typename O>
__global__ void __launch_bounds__(WAVE)
w2_kernel(const R *sim, const O *obs, R *adj, double *mis, int T, double *scratch, int *flags)
{
    extern __shared__ double lds[];
    double *Gs = lds;          // g, then G
    double *Fs = lds + T;      // f, then F
    const int tr = blockIdx.x;
    const int lane = threadIdx.x;
    const R *f0 = sim + (size_t)tr * T;
    const O *g0 = obs + (size_t)tr * T;
    double *h = scratch + (size_t)tr * 3 * T;    // h, then D
    double *sx = h + T, *st = h + 2 * T;          // sorted copies when needed
    const double step = 1.0 / (double)(T - 1);
    const double dt = tval(1, T, step) - tval(0, T, step);

    double mn = INFINITY, mx = -INFINITY;
    for (int k = lane; k < T; k += WAVE) {
        const double v = (double)g0[k];
        mn = fmin(mn, v);
        mx = fmax(mx, v);
    }
    mn = wave_min(mn);
    mx = wave_max(mx);
    const double c = fmax(-1.1 * mn, 0.1 * mx);

    double sg = 0.0, sf = 0.0;
    for (int k = lane; k < T; k += WAVE) {
        const double g = (double)g0[k] + c, f = (double)f0[k] + c;
        Gs[k] = g;
        Fs[k] = f;
        sg += g * dt;
        sf += f * dt;
    }
    sg = wave_sum(sg);
    sf = wave_sum(sf);
    // an all-zero observed trace gives c = 0 and a 0/0 normalization (NaN, as
    // in the reference); report it
    if (!(sg > 0.0) && lane == 0) flags[tr] = 2;
    __syncthreads();
    for (int k = lane; k < T; k += WAVE) {
        Gs[k] = Gs[k] / sg;
        Fs[k] = Fs[k] / sf;
    }
    __syncthreads();
    cumtrapz_inplace<false>(Gs, T, step);
    cumtrapz_inplace<false>(Fs, T, step);

    const bool gmono = is_monotone(Gs, T), fmono = is_monotone(Fs, T);
    const double *gx = Gs, *gy = nullptr, *fx = Fs, *fy = nullptr;
    if (!fmono) {
        rank_sort(Fs, T, step, sx, st);
        fx = sx;
        fy = st;
    }
    // c >= -1.1 min(obs) and c >= 0.1 max(obs) give obs + c >= 0 on every
    // sample, so G never decreases for finite input; the flag reports the
    // non-finite case to the host.
    if (!gmono && lane == 0 && flags[tr] == 0) flags[tr] = 1;

    auto tgrid = [=](int i) { return tval(i, T, step); };
    auto Gx = [=](int i) { return gx[i]; };
    auto Fx = [=](int i) { return fx[i]; };
    auto Fy = [=](int i) { return fy ? fy[i] : tval(i, T, step); };
    (void)gy;

    double w = 0.0;
    for (int k = lane; k < T; k += WAVE) {
        const double tk = tgrid(k);
        const double Fi = interp_at_grid(Fs, k, T, step);
        const double q = interp_lin(Gx, tgrid, T, Fi);
        const double F1 = interp_lin(Fx, Fy, T, tk);
        const double G1 = interp_lin(Gx, tgrid, T, tk);
        h[k] = tk - q;
        const double e = F1 - G1;
        w += e * e;
    }
    w = wave_sum(w);
    __syncthreads();
    // D = 2 * cumtrapz(h) with unit spacing, in the h buffer
    cumtrapz_inplace<true>(h, T, step);
    for (int k = lane; k < T; k += WAVE)
        h[k] = 2.0 * h[k];
    __syncthreads();
    const double D1 = [&] {
        const double xl = tval(T - 2, T, step), xh = tval(T - 1, T, step);
        return (h[T - 1] - h[T - 2]) / (xh - xl) * (1.0 - xl) + h[T - 2];
    }();
    R *a = adj + (size_t)tr * T;
    for (int k = lane; k < T; k += WAVE)
        a[k] = (R)(interp_at_grid(h, k, T, step) - D1);
    if (lane == 0) mis[tr] = w;
    __syncthreads();
}

// In-place exclusive prefix sum over y[0..T) (lane-owned contiguous chunks):
// y[k] <- sum_{i<k} y[i]; y[T] is not touched, the total is returned.
__device__ double excl_prefix_inplace(double *y, int T)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const int C = (T + WAVE - 1) / WAVE;
    const int k0 = lane * C, k1 = min(k0 + C, T);
    double local = 0.0;
    for (int k = k0; k < k1; k++) local += y[k];
    double run = wave_excl_scan(local);
    const double total = wave_sum(local);
    __syncthreads();
    for (int k = k0; k < k1; k++) {
        const double v = y[k];
        y[k] = run;
        run += v;
    }
    __syncthreads();
    return total;
}

// Inclusive running sum in place: y[k] <- sum_{i<=k} y[i] (lane-owned
// contiguous chunks).
__device__ void incl_prefix_inplace(double *y, int T)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const int C = (T + WAVE - 1) / WAVE;
    const int k0 = lane * C, k1 = min(k0 + C, T);
    double local = 0.0;
    for (int k = k0; k < k1; k++) local += y[k];
    double run = wave_excl_scan(local);
    __syncthreads();
    for (int k = k0; k < k1; k++) {
        run += y[k];
        y[k] = run;
    }
    __syncthreads();
}

// Exclusive suffix sum in place: y[k] <- sum_{j>k} y[j].  Lane l owns the
// l-th chunk counted from the end, so the scan over lanes runs from the end.
__device__ void suffix_excl_inplace(double *y, int T)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const int C = (T + WAVE - 1) / WAVE;
    const int r0 = lane * C, r1 = min(r0 + C, T);     // reversed indices r = T-1-k
    double local = 0.0;
    for (int r = r0; r < r1; r++) local += y[T - 1 - r];
    double run = wave_excl_scan(local);
    __syncthreads();
    for (int r = r0; r < r1; r++) {
        const double v = y[T - 1 - r];
        y[T - 1 - r] = run;
        run += v;
    }
    __syncthreads();
}

// The four further Wasserstein-type misfits of adjoint_functions.py, one
// wavefront per trace, all in FP64, t_k = k/(T-1), dt = t_1 - t_0:
//  W2V_QUAD (wasserstein_quad, :70-104): densities sim^2, obs^2 shifted by
//    0.05 max(obs^2); CDFs cumsum(x dt) (rectangle rule, first entry x_0 dt);
//    adjoint D(t) - D(1), D = (2 cumsum(t - G^-1(F(t)))) dt;
//    misfit sum (F^-1(t) - G^-1(t))^2;
//  W2V_EXP (wasserstein_experimental, :107-141): the shift of wasserstein,
//    rectangle-rule CDFs, adjoint Y(t) - Y(1) with
//    Y = t ((t - q)^2 - 2 t q), q = G^-1(F(t)); misfit
//    sum (F^-1(t) - G^-1(t))^2 F^-1(t);
//  W2V_PM (wasserstein_pm, :144-223): densities max(0, x) shifted by
//    0.01 max(max(0, obs)); the reference builds its negative-part CDFs from
//    the positive parts, so both halves are the positive part: adjoint
//    dist(t) - dist(1), dist = cumsum(t - q) dt + cumsum(t - q) dt; misfit
//    twice sum (F^-1(t) - G^-1(t))^2;
//  W2V_YANG (wasserstein_yang, :226-266): shift c = -1.1 min(obs) (no
//    floor), rectangle-rule CDFs, q = G^-1(F(t)) (G stably sorted when it
//    decreases), r = t - q, w = -2 f(t) / g(q) (f, g the normalized
//    densities, g interpolated on the t grid with extrapolation); adjoint
//    s_k = sum_{j>k} w_j r_j + (w_k + r_k) r_k (the reference's
//    (triu(1) diag(w) + diag(r)) r); misfit sum (t - G^-1(F))^2 f.
// Inverse CDFs are interp1d's: a stable sort of the abscissae, searchsorted
// left, clipped brackets, linear extrapolation.  scratch: 6 T doubles per
// trace; flags as w2_kernel (2: the observed density sums to 0, 1: the
// observed CDF of QUAD, EXP or PM decreases).
constexpr int W2V_QUAD = 3, W2V_EXP = 4, W2V_PM = 5, W2V_YANG = 6;

template <int V>
__device__ __forceinline__ double w2v_density(double x)
{
    if constexpr (V == W2V_QUAD) return x * x;
    else if constexpr (V == W2V_PM) return x > 0.0 ? x : 0.0;   // np.maximum(0, x)
    else return x;
}

template <typename R, typename O, int V>
__global__ void __launch_bounds__(WAVE)
w2v_kernel(const R *sim, const O *obs, R *adj, double *mis, int T, double *scratch, int *flags)
{
    extern __shared__ double lds[];
    double *Gs = lds;          // G
    double *Fs = lds + T;      // F
    const int tr = blockIdx.x;
    const int lane = threadIdx.x;
    const R *sv = sim + (size_t)tr * T;
    const O *ov = obs + (size_t)tr * T;
    double *gr = scratch + (size_t)tr * 6 * T;   // normalized obs density g
    double *fr = gr + T;                          // normalized sim density f
    double *sx = gr + 2 * T, *st = gr + 3 * T;    // stably sorted CDF and its times
    double *h = gr + 4 * T;                       // transport term, then the adjoint
    double *dg = gr + 5 * T;                      // YANG: (w_k + r_k) r_k
    const double step = 1.0 / (double)(T - 1);
    const double dt = tval(1, T, step) - tval(0, T, step);

    double mn = INFINITY, mx = -INFINITY;
    for (int k = lane; k < T; k += WAVE) {
        const double o = w2v_density<V>((double)ov[k]);
        mn = fmin(mn, o);
        mx = fmax(mx, o);
    }
    mn = wave_min(mn);
    mx = wave_max(mx);
    double c;
    if constexpr (V == W2V_QUAD) c = mx * 0.05;
    else if constexpr (V == W2V_PM) c = mx * 0.01;
    else if constexpr (V == W2V_EXP) c = fmax(-1.1 * mn, 0.1 * mx);
    else c = -1.1 * mn;
    double sg = 0.0, sf = 0.0;
    for (int k = lane; k < T; k += WAVE) {
        const double g = w2v_density<V>((double)ov[k]) + c;
        const double f = w2v_density<V>((double)sv[k]) + c;
        gr[k] = g;
        fr[k] = f;
        sg += g * dt;
        sf += f * dt;
    }
    sg = wave_sum(sg);
    sf = wave_sum(sf);
    if (!(sg != 0.0) && lane == 0) flags[tr] = 2;
    __syncthreads();
    for (int k = lane; k < T; k += WAVE) {
        const double g = gr[k] / sg, f = fr[k] / sf;
        gr[k] = g;
        fr[k] = f;
        Gs[k] = g * dt;
        Fs[k] = f * dt;
    }
    __syncthreads();
    incl_prefix_inplace(Gs, T);
    incl_prefix_inplace(Fs, T);

    const bool gmono = is_monotone(Gs, T), fmono = is_monotone(Fs, T);
    const double *gx = Gs, *gy = nullptr, *fx = Fs, *fy = nullptr;
    if constexpr (V == W2V_YANG) {
        if (!gmono) {
            rank_sort(Gs, T, step, sx, st);
            gx = sx;
            gy = st;
        }
    } else {
        if (!gmono && lane == 0 && flags[tr] == 0) flags[tr] = 1;
        if (V == W2V_EXP && !fmono) {
            rank_sort(Fs, T, step, sx, st);
            fx = sx;
            fy = st;
        }
    }
    auto tgrid = [=](int i) { return tval(i, T, step); };
    auto Gx = [=](int i) { return gx[i]; };
    auto Gy = [=](int i) { return gy ? gy[i] : tval(i, T, step); };
    auto Fx = [=](int i) { return fx[i]; };
    auto Fy = [=](int i) { return fy ? fy[i] : tval(i, T, step); };
    auto gv = [=](int i) { return gr[i]; };

    double w = 0.0;
    for (int k = lane; k < T; k += WAVE) {
        const double tk = tgrid(k);
        const double q = interp_lin(Gx, Gy, T, interp_at_grid(Fs, k, T, step));
        const double r = tk - q;
        if constexpr (V == W2V_QUAD || V == W2V_PM) {
            h[k] = r;
            const double e = interp_lin(Fx, Fy, T, tk) - interp_lin(Gx, Gy, T, tk);
            w += e * e;
        } else if constexpr (V == W2V_EXP) {
            h[k] = tk * (r * r - 2.0 * tk * q);
            const double F1 = interp_lin(Fx, Fy, T, tk);
            const double e = F1 - interp_lin(Gx, Gy, T, tk);
            w += e * e * F1;
        } else {
            const double fi = interp_at_grid(fr, k, T, step);
            const double wk = -2.0 * fi / interp_lin(tgrid, gv, T, q);
            h[k] = wk * r;
            dg[k] = (wk + r) * r;
            const double e = tk - interp_lin(Gx, Gy, T, Fs[k]);
            w += e * e * fr[k];
        }
    }
    w = wave_sum(w);
    __syncthreads();
    R *a = adj + (size_t)tr * T;
    if constexpr (V == W2V_YANG) {
        suffix_excl_inplace(h, T);
        for (int k = lane; k < T; k += WAVE) a[k] = (R)(h[k] + dg[k]);
    } else {
        if constexpr (V == W2V_QUAD || V == W2V_PM) {
            incl_prefix_inplace(h, T);
            for (int k = lane; k < T; k += WAVE) {
                if constexpr (V == W2V_QUAD) h[k] = 2.0 * h[k] * dt;
                else h[k] = h[k] * dt + h[k] * dt;
            }
            __syncthreads();
        }
        const double xl = tval(T - 2, T, step), xh = tval(T - 1, T, step);
        const double D1 = (h[T - 1] - h[T - 2]) / (xh - xl) * (1.0 - xl) + h[T - 2];
        for (int k = lane; k < T; k += WAVE) a[k] = (R)(interp_at_grid(h, k, T, step) - D1);
    }
    if (lane == 0) mis[tr] = (V == W2V_PM) ? w + w : w;
    __syncthreads();
}

template <typename R, typename O, int V>
hipError_t launch_w2v(const R *sim, const O *obs, R *adj, double *mis, int ntraces, int T,
                      double *scratch, hipStream_t s)
{
    const size_t lds = (size_t)2 * T * sizeof(double);
    hipError_t e = hipFuncSetAttribute((const void *)w2v_kernel<R, O, V>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    int *flags = misfit_flags(V, scratch, ntraces, T);
    hipLaunchKernelGGL((w2v_kernel<R, O, V>), dim3(ntraces), dim3(WAVE), lds, s, sim, obs, adj,
                       mis, T, scratch, flags);
    return hipGetLastError();
}

// Exact derivative of the reference W2 value W = sum_k (F^-1(t_k) - G^-1(t_k))^2
// with respect to the simulated trace, for a monotone F (sim + c >= 0).
// F^-1(t_k) = t_lo + (t_hi - t_lo) (t_k - F_lo) / (F_hi - F_lo) on the
// bracket lo = clip(searchsorted_left(F, t_k), 1, T-1) - 1, hi = lo + 1, so
//   dW/dF_lo += 2 e_k (t_hi - t_lo) (t_k - F_hi) / (F_hi - F_lo)^2 =: A_k,
//   dW/dF_hi += -2 e_k (t_hi - t_lo) (t_k - F_lo) / (F_hi - F_lo)^2 =: B_k.
// lo is non-decreasing in k, so bracket j collects the contiguous range of k
// with t_k in (F_j, F_{j+1}] from prefix sums of A and B (no atomics).  Then
// the transposes of the trapezoid CDF, F_j = sum_{l=1..j} d_l (f_l + f_{l-1})/2,
// and of the normalization f = h / S, h = sim + c, S = sum(h dt).
// flags: 1 G not monotone, 2 all-zero observed trace, 3 F not monotone.
template <typename R, typename O>
__global__ void __launch_bounds__(WAVE)
w2x_kernel(const R *sim, const O *obs, R *rho, double *mis, int T, double *scratch, int *flags)
{
    extern __shared__ double lds[];
    double *Gs = lds;
    double *Fs = lds + T;
    const int tr = blockIdx.x;
    const int lane = threadIdx.x;
    const R *f0 = sim + (size_t)tr * T;
    const O *g0 = obs + (size_t)tr * T;
    double *PA = scratch + (size_t)tr * 3 * T;   // A_k, then its exclusive prefix sums
    double *PB = PA + T;                         // B_k, likewise
    double *Fb = PB + T;                         // dW/dF, then its exclusive prefix sums
    const double step = 1.0 / (double)(T - 1);
    const double dt = tval(1, T, step) - tval(0, T, step);

    double mn = INFINITY, mx = -INFINITY;
    for (int k = lane; k < T; k += WAVE) {
        const double v = (double)g0[k];
        mn = fmin(mn, v);
        mx = fmax(mx, v);
    }
    mn = wave_min(mn);
    mx = wave_max(mx);
    const double c = fmax(-1.1 * mn, 0.1 * mx);
    double sg = 0.0, sf = 0.0;
    for (int k = lane; k < T; k += WAVE) {
        const double g = (double)g0[k] + c, f = (double)f0[k] + c;
        Gs[k] = g;
        Fs[k] = f;
        sg += g * dt;
        sf += f * dt;
    }
    sg = wave_sum(sg);
    sf = wave_sum(sf);
    if (!(sg > 0.0) && lane == 0) flags[tr] = 2;
    __syncthreads();
    for (int k = lane; k < T; k += WAVE) {
        Gs[k] = Gs[k] / sg;
        Fs[k] = Fs[k] / sf;
    }
    __syncthreads();
    cumtrapz_inplace<false>(Gs, T, step);
    cumtrapz_inplace<false>(Fs, T, step);
    const bool gmono = is_monotone(Gs, T), fmono = is_monotone(Fs, T);
    if (lane == 0 && flags[tr] == 0 && (!gmono || !fmono)) flags[tr] = gmono ? 3 : 1;
    auto tgrid = [=](int i) { return tval(i, T, step); };
    auto Gx = [=](int i) { return Gs[i]; };
    auto Fx = [=](int i) { return Fs[i]; };

    double w = 0.0;
    for (int k = lane; k < T; k += WAVE) {
        const double tk = tgrid(k);
        int i = lower_bound(Fx, T, tk);
        i = i < 1 ? 1 : (i > T - 1 ? T - 1 : i);
        const double Fl = Fs[i - 1], Fh = Fs[i], tl = tgrid(i - 1), th = tgrid(i);
        const double F1 = (th - tl) / (Fh - Fl) * (tk - Fl) + tl;
        const double G1 = interp_lin(Gx, tgrid, T, tk);
        const double e = F1 - G1;
        w += e * e;
        const double d2 = (Fh - Fl) * (Fh - Fl);
        PA[k] = 2.0 * e * (th - tl) * (tk - Fh) / d2;
        PB[k] = -2.0 * e * (th - tl) * (tk - Fl) / d2;
    }
    w = wave_sum(w);
    __syncthreads();
    const double totA = excl_prefix_inplace(PA, T);
    const double totB = excl_prefix_inplace(PB, T);
    // ks(j) = first k with t_k > F_j (bracket of k >= j); ks(0) = 0, ks(T-1) = T
    auto ks = [&](int j) -> int {
        if (j <= 0) return 0;
        if (j >= T - 1) return T;
        const double Fj = Fs[j];
        int lo = 0, hi = T;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (Fj < tgrid(mid)) hi = mid; else lo = mid + 1;
        }
        return lo;
    };
    auto preA = [&](int k) { return k < T ? PA[k] : totA; };
    auto preB = [&](int k) { return k < T ? PB[k] : totB; };
    for (int j = lane; j < T; j += WAVE) {
        double v = 0.0;
        if (j <= T - 2) v += preA(ks(j + 1)) - preA(ks(j));
        if (j >= 1) v += preB(ks(j)) - preB(ks(j - 1));
        Fb[j] = v;
    }
    __syncthreads();
    const double totF = excl_prefix_inplace(Fb, T);   // Fb[j] = sum_{i<j}; suffix S_j = totF - Fb[j]
    auto fbar = [&](int m) -> double {
        double v = 0.0;
        if (m >= 1) v += 0.5 * (tgrid(m) - tgrid(m - 1)) * (totF - Fb[m]);
        if (m + 1 <= T - 1) v += 0.5 * (tgrid(m + 1) - tgrid(m)) * (totF - Fb[m + 1]);
        return v;
    };
    double dot = 0.0;
    for (int m = lane; m < T; m += WAVE) dot += fbar(m) * ((double)f0[m] + c);
    dot = wave_sum(dot);
    R *o = rho + (size_t)tr * T;
    for (int m = lane; m < T; m += WAVE) o[m] = (R)(fbar(m) / sf - dt * dot / (sf * sf));
    if (lane == 0) mis[tr] = w;
}

// Exact L2: misfit sum (sim - obs)^2, derivative 2 (sim - obs).
template <typename R, typename O>
__global__ void __launch_bounds__(WAVE)
l2x_kernel(const R *sim, const O *obs, R *rho, double *mis, int T)
{
    const int tr = blockIdx.x, lane = threadIdx.x;
    double w = 0.0;
    for (int k = lane; k < T; k += WAVE) {
        const double r = (double)sim[(size_t)tr * T + k] - (double)obs[(size_t)tr * T + k];
        rho[(size_t)tr * T + k] = (R)(2.0 * r);
        w += r * r;
    }
    w = wave_sum(w);
    if (lane == 0) mis[tr] = w;
}

template <typename R, typename O>
__global__ void __launch_bounds__(WAVE)
l2_kernel(const R *sim, const O *obs, R *adj, double *mis, int T)
{
    const int tr = blockIdx.x, lane = threadIdx.x;
    double w = 0.0;
    for (int k = lane; k < T; k += WAVE) {
        const double r = (double)sim[(size_t)tr * T + k] - (double)obs[(size_t)tr * T + k];
        adj[(size_t)tr * T + k] = (R)r;
        w += r * r;
    }
    w = wave_sum(w);
    if (lane == 0) mis[tr] = w;
}

// ---------------------------------------------------------------------------
// Squared-envelope misfit (adjoint_functions.quad_envelope,
// adjoint_functions.py:14-22):
//   Hs = hilbert(sim), Ho = hilbert(obs) (analytic signals, per trace);
//   E = |Hs|^2 - |Ho|^2; misfit = sum E^2;
//   adj = sim E - Im(hilbert(E Im(Hs))).
// scipy.signal.hilbert is ifft(fft(x) h) with h = 1 at k = 0 (and k = N/2 for
// even N), 2 for 0 < k < N/2, 0 above.  Its real part is x and its imaginary
// part is the circular convolution y[n] = sum_m c[(n - m) mod N] x[m] with
// c[j] = (2/N) sum_{k=1}^{K} sin(2 pi k j / N), K = ceil(N/2) - 1
// (hilbert_table).  Here y is formed in FP64 as a tiled product of the
// traces with the circulant matrix, without an FFT.
// ---------------------------------------------------------------------------
void hilbert_table(int N, double *c)
{
    // sum_{k=1}^{K} sin(k t) = sin(K t / 2) sin((K + 1) t / 2) / sin(t / 2),
    // t = 2 pi j / N; the arguments are reduced modulo 2 pi in integers
    const int K = (N + 1) / 2 - 1;
    const double pi = 3.14159265358979323846;
    auto sinpi = [&](long long num) {   // sin(pi num / N)
        const long long r = ((num % (2LL * N)) + 2LL * N) % (2LL * N);
        return std::sin(pi * (double)r / (double)N);
    };
    c[0] = 0.0;
    for (int j = 1; j < N; j++) {
        const double v = sinpi((long long)K * j) * sinpi((long long)(K + 1) * j) / sinpi(j);
        c[j] = 2.0 * v / (double)N;
    }
}

// y[tr][n] = sum_m c[(n - m) mod N] x[tr][m] for a 64-trace x 64-sample tile
// per work-group: the K dimension (m) in chunks of 32 through LDS, the 95
// circulant values a chunk needs staged beside it; each thread keeps a 4 x 4
// block of outputs (traces ty + 16 i, samples tx + 16 j).
constexpr int HB = 64, HK = 32;
template <typename I>
__global__ void __launch_bounds__(256)
hilbert_kernel(const I *x, double *y, const double *c, int ntr, int N)
{
    __shared__ double xs[HB][HK + 1];
    __shared__ double cs[HB + HK];
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    const int n0 = blockIdx.x * HB, t0 = blockIdx.y * HB;
    double acc[4][4] = {};
    for (int m0 = 0; m0 < N; m0 += HK) {
        for (int v = tid; v < HB * HK; v += 256) {
            const int r = v / HK, k = v - r * HK;
            const int tr = t0 + r, m = m0 + k;
            xs[r][k] = (tr < ntr && m < N) ? (double)x[(size_t)tr * N + m] : 0.0;
        }
        // cs[d] = c[(n0 - m0 - (HK - 1) + d) mod N], d = 0 .. HB + HK - 2
        for (int d = tid; d < HB + HK - 1; d += 256) {
            int j = (n0 - m0 - (HK - 1) + d) % N;
            if (j < 0) j += N;
            cs[d] = c[j];
        }
        __syncthreads();
#pragma unroll 8
        for (int k = 0; k < HK; k++) {
            double xv[4], cv[4];
#pragma unroll
            for (int i = 0; i < 4; i++) xv[i] = xs[ty + 16 * i][k];
#pragma unroll
            for (int j = 0; j < 4; j++) cv[j] = cs[tx + 16 * j - k + (HK - 1)];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] = fma(cv[j], xv[i], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int tr = t0 + ty + 16 * i;
        if (tr >= ntr) continue;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int n = n0 + tx + 16 * j;
            if (n < N) y[(size_t)tr * N + n] = acc[i][j];
        }
    }
}

// E = |Hs|^2 - |Ho|^2 into e, q = E Im(Hs) into q (may alias hs / ho: each
// element is read before it is written), misfit of the trace = sum E^2.
template <typename R, typename O>
__global__ void __launch_bounds__(WAVE)
env_mid_kernel(const R *sim, const O *obs, const double *hs, const double *ho, double *e,
               double *q, double *mis, int T)
{
    const int tr = blockIdx.x, lane = threadIdx.x;
    double w = 0.0;
    for (int k = lane; k < T; k += WAVE) {
        const size_t i = (size_t)tr * T + k;
        const double s = (double)sim[i], o = (double)obs[i], a = hs[i], b = ho[i];
        const double E = (s * s + a * a) - (o * o + b * b);
        e[i] = E;
        q[i] = E * a;
        w += E * E;
    }
    w = wave_sum(w);
    if (lane == 0) mis[tr] = w;
}

// adj = sim E - Im(hilbert(E Im(Hs)))
template <typename R>
__global__ void env_adj_kernel(const R *sim, const double *e, const double *hq, R *adj, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) adj[i] = (R)((double)sim[i] * e[i] - hq[i]);
}

template <typename I>
hipError_t launch_hilbert(const I *x, double *y, const double *c, int ntr, int N, hipStream_t s)
{
    const dim3 grid((N + HB - 1) / HB, (ntr + HB - 1) / HB);
    hipLaunchKernelGGL((hilbert_kernel<I>), grid, dim3(256), 0, s, x, y, c, ntr, N);
    return hipGetLastError();
}

// Envelope misfit of ntraces traces: scratch holds 3 ntraces T doubles.
template <typename R, typename O>
hipError_t launch_envelope(const R *sim, const O *obs, R *adj, double *mis, int ntraces, int T,
                           double *scratch, const double *hilb, hipStream_t s)
{
    if (!hilb) return hipErrorInvalidValue;
    const size_t n = (size_t)ntraces * T;
    double *hs = scratch, *ho = scratch + n, *hq = scratch + 2 * n;
    hipError_t e = launch_hilbert<R>(sim, hs, hilb, ntraces, T, s);
    if (e == hipSuccess) e = launch_hilbert<O>(obs, ho, hilb, ntraces, T, s);
    if (e != hipSuccess) return e;
    // E over hs, q over ho
    hipLaunchKernelGGL((env_mid_kernel<R, O>), dim3(ntraces), dim3(WAVE), 0, s, sim, obs, hs, ho,
                       hs, ho, mis, T);
    e = hipGetLastError();
    if (e == hipSuccess) e = launch_hilbert<double>(ho, hq, hilb, ntraces, T, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((env_adj_kernel<R>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, sim,
                       hs, hq, adj, n);
    return hipGetLastError();
}

// jcum[t] = sum_{tau=1..t} adj[T-1-tau] (jcum[0] = 0): the running integral of
// the flipped adjoint source (somaFonteIntegral with flip=1,
// regression_PML_single.cu:276-292, 803), summed sequentially in the state
// precision as the reference's integral[n] += source[...] does: one thread
// per trace, so every J[k] rounds exactly as the reference's running sum.
template <typename R>
__global__ void __launch_bounds__(WAVE)
jcum_kernel(const R *adj, R *jcum, int ntraces, int T)
{
    const int tr = blockIdx.x * WAVE + threadIdx.x;
    if (tr >= ntraces) return;
    const R *a = adj + (size_t)tr * T;
    R *J = jcum + (size_t)tr * T;
    R run = R(0);
    J[0] = run;
    for (int k = 1; k < T; k++) {
        run = run + a[T - 1 - k];
        J[k] = run;
    }
}

template <typename R>
hipError_t launch_jcum(const R *adj, R *jcum, int ntraces, int T, hipStream_t s)
{
    if (ntraces <= 0) return hipSuccess;
    hipLaunchKernelGGL((jcum_kernel<R>), dim3((ntraces + WAVE - 1) / WAVE), dim3(WAVE), 0, s,
                       adj, jcum, ntraces, T);
    return hipGetLastError();
}
template hipError_t launch_jcum<float>(const float *, float *, int, int, hipStream_t);
template hipError_t launch_jcum<double>(const double *, double *, int, int, hipStream_t);

template <typename R, typename O>
hipError_t launch_misfit(int kind, const R *sim, const O *obs, R *adj, R *jcum, double *mis,
                         int ntraces, int T, double *scratch, hipStream_t s, const double *hilb)
{
    if (ntraces <= 0) return hipSuccess;
    if (kind == 0) {
        hipLaunchKernelGGL((l2_kernel<R, O>), dim3(ntraces), dim3(WAVE), 0, s, sim, obs, adj, mis,
                           T);
    } else if (kind == 2) {
        const hipError_t e = launch_envelope<R, O>(sim, obs, adj, mis, ntraces, T, scratch, hilb, s);
        if (e != hipSuccess) return e;
    } else if (kind >= W2V_QUAD && kind <= W2V_YANG) {
        hipError_t e;
        switch (kind) {
        case W2V_QUAD: e = launch_w2v<R, O, W2V_QUAD>(sim, obs, adj, mis, ntraces, T, scratch, s); break;
        case W2V_EXP: e = launch_w2v<R, O, W2V_EXP>(sim, obs, adj, mis, ntraces, T, scratch, s); break;
        case W2V_PM: e = launch_w2v<R, O, W2V_PM>(sim, obs, adj, mis, ntraces, T, scratch, s); break;
        default: e = launch_w2v<R, O, W2V_YANG>(sim, obs, adj, mis, ntraces, T, scratch, s); break;
        }
        if (e != hipSuccess) return e;
    } else {
        const size_t lds = (size_t)2 * T * sizeof(double);
        hipError_t e = hipFuncSetAttribute((const void *)w2_kernel<R, O>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        int *flags = misfit_flags(kind, scratch, ntraces, T);
        hipLaunchKernelGGL((w2_kernel<R, O>), dim3(ntraces), dim3(WAVE), lds, s, sim, obs, adj,
                           mis, T, scratch, flags);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !jcum) return e;
    return launch_jcum<R>(adj, jcum, ntraces, T, s);
}
template hipError_t launch_misfit<float, float>(int, const float *, const float *, float *,
                                                float *, double *, int, int, double *,
                                                hipStream_t, const double *);
template hipError_t launch_misfit<double, float>(int, const double *, const float *, double *,
                                                 double *, double *, int, int, double *,
                                                 hipStream_t, const double *);
template hipError_t launch_misfit<double, double>(int, const double *, const double *, double *,
                                                  double *, double *, int, int, double *,
                                                  hipStream_t, const double *);

template <typename R, typename O>
hipError_t launch_misfit_exact(int kind, const R *sim, const O *obs, R *rho, double *mis,
                               int ntraces, int T, double *scratch, int *flags, hipStream_t s)
{
    if (ntraces <= 0) return hipSuccess;
    if (kind == 0) {
        hipLaunchKernelGGL((l2x_kernel<R, O>), dim3(ntraces), dim3(WAVE), 0, s, sim, obs, rho, mis,
                           T);
    } else {
        const size_t lds = (size_t)2 * T * sizeof(double);
        hipError_t e = hipFuncSetAttribute((const void *)w2x_kernel<R, O>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((w2x_kernel<R, O>), dim3(ntraces), dim3(WAVE), lds, s, sim, obs, rho,
                           mis, T, scratch, flags);
    }
    return hipGetLastError();
}
template hipError_t launch_misfit_exact<float, float>(int, const float *, const float *, float *,
                                                      double *, int, int, double *, int *,
                                                      hipStream_t);
template hipError_t launch_misfit_exact<double, float>(int, const double *, const float *,
                                                       double *, double *, int, int, double *,
                                                       int *, hipStream_t);
template hipError_t launch_misfit_exact<double, double>(int, const double *, const double *,
                                                        double *, double *, int, int, double *,
                                                        int *, hipStream_t);

template <typename R>
__global__ void accum_grad_kernel(const R *grad, int nshots, int X, int Z, int ld,
                                  long long sstride, double *acc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= X * Z) return;
    const int z = i / X, x = i - z * X;
    const R *g = grad + (size_t)(z + ZOFF) * ld + x + XOFF;
    double a = 0.0;
    for (int s = 0; s < nshots; s++)
        a += (double)g[(size_t)s * sstride];
    acc[i] += a;
}

template <typename R>
hipError_t launch_accum_grad(const R *grad, int nshots, int X, int Z, int ld, long long sstride,
                             double *acc, hipStream_t s)
{
    const int n = X * Z;
    hipLaunchKernelGGL((accum_grad_kernel<R>), dim3((n + 255) / 256), dim3(256), 0, s, grad,
                       nshots, X, Z, ld, sstride, acc);
    return hipGetLastError();
}
template hipError_t launch_accum_grad<float>(const float *, int, int, int, int, long long,
                                             double *, hipStream_t);
template hipError_t launch_accum_grad<double>(const double *, int, int, int, int, long long,
                                              double *, hipStream_t);

__global__ void finalize_grad_kernel(const double *acc, int n, double scale, float *o32)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) o32[i] = (float)(acc[i] * scale);
}

hipError_t launch_finalize_grad(const double *acc, int n, double scale, float *o32, hipStream_t s)
{
    hipLaunchKernelGGL(finalize_grad_kernel, dim3((n + 255) / 256), dim3(256), 0, s, acc, n, scale,
                       o32);
    return hipGetLastError();
}

// Model array of the step kernels from c^2 on the device: cm at (z + ZOFF) *
// ld + x + XOFF = (c^2 dt) / dx in R, the padding left as it is (zero), and
// *bad lowered to 1 + the smallest flat index whose c^2 is not finite or is
// negative (the caller sets it above any such value first).  One product and
// one correctly rounded division, as the host forms it.
template <typename R>
__global__ void model_cm_kernel(const float *__restrict__ cq, int X, int Z, int ld, R dt, R dx,
                                R *__restrict__ cm, int *__restrict__ bad)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y;
    if (x >= X || z >= Z) return;
    const size_t q = (size_t)z * X + x;
    const float c = cq[q];
    if (!(c >= 0.0f && c <= 3.402823466e38f)) atomicMin(bad, (int)q + 1);
    cm[(size_t)(z + ZOFF) * ld + x + XOFF] = ((R)c * dt) / dx;
}

template <typename R>
hipError_t launch_model_cm(const float *cq, int X, int Z, int ld, R dt, R dx, R *cm, int *bad,
                           hipStream_t s)
{
    hipLaunchKernelGGL((model_cm_kernel<R>), dim3((X + 255) / 256, Z), dim3(256), 0, s, cq, X, Z,
                       ld, dt, dx, cm, bad);
    return hipGetLastError();
}
template hipError_t launch_model_cm<float>(const float *, int, int, int, float, float, float *,
                                           int *, hipStream_t);
template hipError_t launch_model_cm<double>(const float *, int, int, int, double, double,
                                            double *, int *, hipStream_t);

// Source-side illumination: ill[z*X+x] += w * sum_s P_s(t)[z][x]^2 over the
// shots of the batch (summed in shot order, FP32), one sampled step t.
template <typename R>
__global__ void illum_kernel(const R *__restrict__ P, int nshots, int X, int Z, int ld,
                             long long sstride, float w, float *__restrict__ ill)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y;
    if (x >= X || z >= Z) return;
    const R *p = P + (size_t)(z + ZOFF) * ld + x + XOFF;
    float a = 0.0f;
    for (int s = 0; s < nshots; s++) {
        const float v = (float)p[(size_t)s * sstride];
        a = fmaf(v, v, a);
    }
    float *o = ill + (size_t)z * X + x;
    *o = fmaf(a, w, *o);
}

template <typename R>
hipError_t launch_illum(const R *P, int nshots, int X, int Z, int ld, long long sstride, float w,
                        float *ill, hipStream_t s)
{
    hipLaunchKernelGGL((illum_kernel<R>), dim3((X + 255) / 256, Z), dim3(256), 0, s, P, nshots, X,
                       Z, ld, sstride, w, ill);
    return hipGetLastError();
}
template hipError_t launch_illum<float>(const float *, int, int, int, int, long long, float,
                                        float *, hipStream_t);
template hipError_t launch_illum<double>(const double *, int, int, int, int, long long, float,
                                         float *, hipStream_t);

__global__ void reduce_misfit_kernel(const double *mis, int nshots, int n_rx, double *out)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nshots) return;
    double acc = 0.0;
    for (int r = 0; r < n_rx; r++)
        acc += mis[(size_t)s * n_rx + r];
    out[s] = acc;
}

hipError_t launch_reduce_misfit(const double *mis, int nshots, int n_rx, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(reduce_misfit_kernel, dim3((nshots + 63) / 64), dim3(64), 0, s, mis, nshots,
                       n_rx, out);
    return hipGetLastError();
}

}  // namespace otfwi
