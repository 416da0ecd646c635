// Internal declarations shared by the HIP kernels and the host orchestration
// of libotfwi_pml.so.  Not part of the public C ABI (include/otfwi_pml.h).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace otfwi {

// Tile geometry of the fused step kernels.  A work-group owns a TW x TH tile
// of cells and recomputes the P half-step on the tile plus a 4-cell halo
// ("E box", EW x (TH+8)) so that the A half-step of the same time step needs
// no exchange with neighbouring tiles.
constexpr int TW = 64;                  // owned columns
constexpr int EW = TW + 8;              // E box columns  x0-4 .. x0+67
constexpr int AXW = TW + 16;            // staged Ax columns x0-8 .. x0+71
constexpr int NTHREADS = 256;           // 16 column quads x 16 row groups
constexpr int BOX = 4;                  // never-updated outer ring width

// Padded field layout.  Every field of every shot is a (ZP x LD) plane with
// the cell (x, z) at (z + ZOFF) * LD + x + XOFF; the padding and every cell
// outside the updated box hold 0 at all times, so stencil loads need neither
// bounds checks nor masks.  XOFF = 64 keeps each tile row 256-byte aligned.
constexpr int XOFF = 64, ZOFF = 8;
inline int round_up(int a, int b) { return (a + b - 1) / b * b; }
inline int padded_ld(int X) { return round_up(X, 64) + 2 * XOFF; }
inline int padded_rows(int Z) { return round_up(Z, 64) + 2 * ZOFF; }

// Tile flags (per tile of a model, for one tile height).
enum : uint8_t {
    TF_STRIP = 1,   // E box touches a damped (split-field) cell
    TF_EVENT = 2,   // E box holds a source, receiver or ring point
    TF_EDGE = 4,    // E box reaches outside the updated box (results need masking)
};

// Point table entry: what happens at a cell flagged in the event map.
struct PointEvt {
    int ring;   // slot in the boundary ring, or -1
    int rx;     // receiver index, or -1
    int src;    // source index, or -1
    int pad;
};

template <typename R>
struct FwdArgs {
    int X, Z, ld, ntx, ntz, nshots, t, T;
    long long sstride;            // elements between shots in a field (ZP*ld)
    const R *P0, *Ax0, *Az0, *Px0;  // state at t-1 (padded planes)
    R *P1, *Ax1, *Az1, *Px1;        // state at t
    const R *cm;                  // (c^2*dt)/dx per cell, padded plane, 0 outside the grid
    R cA;                         // dt/dx
    const R *kx, *kz;             // 1 - d*dt at x + XOFF ([ld]) and z + ZOFF ([ZP]), shifted by one cell
    const uint8_t *tflags;        // [ntz*ntx]
    const int *evt;               // padded plane: index into pts, or -1
    const PointEvt *pts;
    const int *shot_src;          // [nshots] active source (-1: all)
    const R *icum;                // [n_src][T] running sum of each source wavelet
    R *rec;                       // [nshots][n_rx][T] (may be null)
    int n_rx;
    R *ring;                      // [nshots][T][n_ring] (may be null)
    int n_ring;
};

template <typename R>
struct AdjArgs {
    int X, Z, ld, ntx, ntz, nshots, t, T;
    long long sstride;
    // adjoint (lambda) field, damped forward operator
    const R *LP0, *LAx0, *LAz0, *LPx0;
    R *LP1, *LAx1, *LAz1, *LPx1;
    // reconstructed forward field (undamped reverse operator)
    const R *U0, *UAx0, *UAz0;
    R *U1, *UAx1, *UAz1;
    R *grad;                      // [nshots][Z][ld] accumulated imaging condition
    const R *cm;
    R cA;
    const R *kx, *kz;
    const uint8_t *tflags;
    const int *evt;
    const PointEvt *pts;
    const R *jcum;                // [nshots][n_rx][T] running sum of the flipped adjoint source
    int n_rx;
    const R *ring;                // [nshots][T][n_ring]
    int n_ring;
};

// One step t of the exact discrete adjoint (exact.hip): u_t, w_t from
// u_{t+1}, w_{t+1}, a_t, b_t and the forward A(t-1); then a_{t-1}, b_{t-1}.
template <typename R>
struct XAdjArgs {
    int X, Z, ld, nshots, n_rx, T, t;
    long long sstride;
    const R *u0, *w0;             // u_{t+1}, w_{t+1}
    const R *a, *b;               // a_t, b_t
    R *u1, *w1;                   // u_t, w_t
    R *a1, *b1;                   // a_{t-1}, b_{t-1}
    const R *Ax, *Az;             // forward A(t-1), per shot
    R *grad;                      // dJ/dm, per shot
    const R *cm, *kx, *kz;
    R cA;
    const int *rxmap;             // padded plane: receiver index or -1
    const R *rho;                 // [nshots][n_rx][T] dJ/d(trace)
};
template <typename R>
hipError_t launch_xadj_step(const XAdjArgs<R> &a, hipStream_t s);

// Launchers (prop.hip).  th = tile height (16, 32 or 64); persistent_wgs > 0
// launches that many persistent work-groups instead of one per tile.
template <typename R>
hipError_t launch_fwd_step(const FwdArgs<R> &a, int th, int persistent_wgs, hipStream_t s);
template <typename R>
hipError_t launch_adj_step(const AdjArgs<R> &a, int th, int persistent_wgs, hipStream_t s);
// Two time steps t, t+1 per launch (z-marching strips, prop.hip): reads the
// state before the pair from P0/Ax0/Az0/Px0 and writes the state after it to
// P1/Ax1/Az1/Px1 (a different slot).  A work-group of nw = 256 or 384 threads
// owns nw - 32 columns starting at bx * (nw - 32) and computes F2_HALO more
// on each side.
//   mode 0 forward field (fwd_step_kernel x2): samples rec / ring at t, t+1;
//   mode 1 adjoint field lambda (adj_step_kernel's lambda x2): injects jcum,
//          grad += lambda_P(t) lap0 + lambda_P(t+1) lap1;
//   mode 2 reconstructed field U (adj_step_kernel's U x2): ring values
//          written back at ring times T-1-t and T-2-t, writes lap0 / lap1 =
//          (lap9_x + lap9_z)(U(t)) / 5040 and the same of U(t+1); no psi.
// evrow (modes 0, 1): evrow[bx * Z + z] != 0 where row z holds a source or
// receiver in columns [bx * own - F2_HALO, bx * own - F2_HALO + nw), and
// pmap is the padded-plane map of int4 {ring + 1, rx + 1, src + 1, 0}
// (0: none).
// mask (modes 0, 2): f2_ring_words(nw) words per (bx, z) at
// mask[(bx * Z + z) * f2_ring_words(nw)]: the ring slot of the first ring
// cell in the block's columns, the number of ring cells before each
// 32-column word (16 bits each, nw / 64 words), and the nw / 32 lane masks of
// the ring cells; ring slots are numbered in raster order (z, then x).
constexpr int F2_HALO = 16;
inline int f2_own(int nw)
{
    return nw == 128 || nw == 192 || nw == 256 || nw == 384 ? nw - 2 * F2_HALO : 0;
}
inline int f2_ring_words(int nw) { return 1 + nw / 64 + nw / 32; }
// Automatic block width: 256.  384-thread blocks waste fewer threads on
// partial column blocks (1152 vs 1280 threads over 1024 columns) but ran
// 1.3-1.45x slower per launch on MI355X at 64 x 1024^2: two 6-wave work-groups
// per CU hide latency worse than three 4-wave ones.
#ifndef F2_DEFAULT_NW
#define F2_DEFAULT_NW 256
#endif
inline int f2_width(int X) { (void)X; return F2_DEFAULT_NW; }
// The two slots of each field are one allocation: P1 - P0 = +-slot_elems
// (likewise Ax, Az, Px), so one buffer resource per field covers both.
template <typename R>
struct Pair2Args {
    int X, Z, ld, nshots, t, T, n_rx, n_ring;
    int nw;                       // threads per work-group (256 or 384); the tables match it
    long long sstride, slot_elems;
    const R *P0, *Ax0, *Az0, *Px0;
    R *P1, *Ax1, *Az1, *Px1;
    const R *cm, *kx, *kz;
    R cA;
    const int *shot_src;          // mode 0
    const R *icum;                // mode 0
    R *rec;                       // mode 0 (may be null)
    R *ring;                      // mode 0: [nshots][T][n_ring] written (may be null)
    const R *ring_in;             // mode 2: the same, read
    const R *jcum;                // mode 1
    R *grad;                      // mode 1
    R *lap0, *lap1;               // mode 1 reads, mode 2 writes; lap1 = lap0 + slot_elems
    const unsigned *mask;
    const uint8_t *evrow;
    const int *pmap;
};
bool pair2_supported(long long sstride, long long slot_elems, int prec);
// Work-groups of one two-step launch over nshots shots.
long long pair2_groups(int X, int Z, int nw, int nshots);
template <typename R>
hipError_t launch_pair2(const Pair2Args<R> &a, int mode, hipStream_t s);
// Host-side checks launch_pair2 applies to one problem's arguments.
template <typename R>
bool pair2_args_ok(const Pair2Args<R> &a, int mode);
// One FP32 two-step launch of F2_MULTI_NW-thread work-groups over nprob
// problems: tab[nprob] their arguments (each passing pair2_args_ok, with
// nw == F2_MULTI_NW), wg0[nprob + 1] the prefix sums of their work-group
// counts (pair2_groups), nwg = wg0[nprob]; both tables in device memory.
constexpr int F2_MULTI_NW = 256;
hipError_t launch_pair2_multi(const Pair2Args<float> *tab, const int *wg0, int nprob, long long nwg,
                              int mode, hipStream_t s);
// Resident work-groups per CU of a step kernel (0 if the variant does not exist).
int step_occupancy(bool adj, int prec, int th, bool persistent);

// Misfit / reduction launchers (misfit.hip).
// kind: 0 = L2, 1 = W2, 2 = squared envelope, 3-6 = wasserstein_quad,
// _experimental, _pm, _yang.  sim/obs/adj/jcum: [nshots][n_rx][T]; mis:
// [nshots*n_rx] FP64 per trace.  jcum may be null (no running sum wanted).  O is the
// observed-trace type.  scratch: misfit_scratch_per_sample(kind) ntraces T
// doubles, then ntraces ints of flags (misfit_flags).
// hilb: hilbert_table(T) on the device (kind 2 only).
// Per-trace flags of the W2-type kernels inside a misfit scratch area: after
// 3 T doubles per trace (W2, envelope) or 6 T (the four further
// Wasserstein-type misfits, kinds 3-6).
inline int misfit_scratch_per_sample(int kind) { return kind >= 3 ? 6 : 3; }
inline int *misfit_flags(int kind, double *scratch, int ntraces, int T)
{
    return reinterpret_cast<int *>(scratch + (size_t)ntraces * misfit_scratch_per_sample(kind) * T);
}
template <typename R, typename O>
hipError_t launch_misfit(int kind, const R *sim, const O *obs, R *adj, R *jcum, double *mis,
                         int ntraces, int T, double *scratch, hipStream_t s,
                         const double *hilb = nullptr);
// c[0..N): the circular-convolution kernel whose product with a trace is the
// imaginary part of scipy.signal.hilbert's analytic signal (host, FP64).
void hilbert_table(int N, double *c);
// Exact derivative of the misfit value w.r.t. the simulated traces (rho):
// kind 0: 2 (sim - obs); kind 1: derivative of the reference W2 value
// (needs a monotone CDF of sim + c; flags as launch_misfit,
// plus 3 = CDF of the simulated trace not monotone).
template <typename R, typename O>
hipError_t launch_misfit_exact(int kind, const R *sim, const O *obs, R *rho, double *mis,
                               int ntraces, int T, double *scratch, int *flags, hipStream_t s);
// jcum of host-provided adjoint sources.
template <typename R>
hipError_t launch_jcum(const R *adj, R *jcum, int ntraces, int T, hipStream_t s);
// acc[z*X+x] += sum_s grad[s][z][x] (FP64), and out = (float)(acc*scale).
template <typename R>
hipError_t launch_accum_grad(const R *grad, int nshots, int X, int Z, int ld, long long sstride,
                             double *acc, hipStream_t s);
hipError_t launch_finalize_grad(const double *acc, int n, double scale, float *o32, hipStream_t s);
// cm = (c^2 dt) / dx of every grid cell into the padded plane (padding
// untouched); *bad lowered to 1 + the smallest flat index of a c^2 that is
// not finite or is negative, left as it is when there is none.
template <typename R>
hipError_t launch_model_cm(const float *cq, int X, int Z, int ld, R dt, R dx, R *cm, int *bad,
                           hipStream_t s);
// ill[z*X+x] += w * sum_s P[s][z][x]^2 (padded planes P, unpadded FP32 ill).
template <typename R>
hipError_t launch_illum(const R *P, int nshots, int X, int Z, int ld, long long sstride, float w,
                        float *ill, hipStream_t s);
// Per-shot misfit sum over its traces in trace order (FP64).
hipError_t launch_reduce_misfit(const double *mis, int nshots, int n_rx, double *out, hipStream_t s);

}  // namespace otfwi
