// Host orchestration and C ABIs of libotfwi_pml.so (see include/otfwi_pml.h).
//
// A model owns its device buffers and one HIP stream.  A forward run of a
// batch of shots is T-1 launches of the fused forward step (every shot of
// the batch in each launch); a gradient adds the misfit kernels and T-1
// launches of the fused adjoint step, then an FP64 reduction over shots.
// The host loop only enqueues work; nothing synchronizes inside it.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <climits>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "otfwi_internal.h"
#include "otfwi_pml.h"

// Host-side arithmetic that must round like the reference (1 - d*dt,
// (c^2*dt)/dx, running source sums) stays unfused.
#pragma clang fp contract(off)

using namespace otfwi;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCK(expr)                                                                  \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess)                                                        \
            return fail(OTFWI_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                         \
    } while (0)

#define RCK(expr)                        \
    do {                                 \
        int r_ = (expr);                 \
        if (r_ != OTFWI_OK) return r_;   \
    } while (0)

template <typename T>
int dalloc(T **p, size_t n)
{
    *p = nullptr;
    if (n == 0) return OTFWI_OK;
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(OTFWI_E_MEM, "hipMalloc of %zu bytes failed: %s", n * sizeof(T),
                    hipGetErrorString(e));
    }
    return OTFWI_OK;
}

template <typename T>
void dfree(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

}  // namespace

struct Batch {
    int cap = 0;
    void *P[2] = {}, *Ax[2] = {}, *Az[2] = {}, *Px[2] = {};
    void *LP[2] = {}, *LAx[2] = {}, *LAz[2] = {}, *LPx[2] = {};
    void *grad = nullptr, *ring = nullptr, *rec = nullptr, *adj = nullptr, *jcum = nullptr;
    void *lap[2] = {};      // two-step adjoint: lap of the reconstructed field at t, t+1
    int lap_cap = 0;
    double *mis_tr = nullptr, *mis_shot = nullptr, *scratch = nullptr;
    int *shot_src = nullptr;
    float *obs = nullptr;   // obs of the shots in this batch
    size_t ring_elems = 0;
    // which observed traces `obs` holds: shots [obs_s0, obs_s0 + obs_nb) of
    // version obs_ver of the model's observed data (obs_nb = 0: none)
    int obs_s0 = 0, obs_nb = 0;
    unsigned long long obs_ver = 0;
};

struct otfwi_model {
    int X = 0, Z = 0, T = 0, ld = 0, zp = 0, prec = 4;   // ld, zp: padded plane shape
    float dx = 1, dz = 1, dt = 1;
    hipStream_t stream = nullptr;
    void *cm = nullptr, *kx = nullptr, *kz = nullptr, *icum = nullptr;
    float *cq32 = nullptr;          // c^2 as given (FP32)
    bool have_cq = false, have_damp = false;
    std::vector<float> dxs_h, dzs_h;  // shifted damping profiles d_x[.][x+1], d_z[z+1][.]
    std::vector<float> kx_h, kz_h;    // 1 where the cell is undamped along that axis, else 0
    std::vector<int> sx, sz, rxx, rxz, vx, vz;
    std::vector<float> src_h;
    int n_src = 0, n_rx = 0, n_ring = 0;
    int *evt = nullptr;
    PointEvt *pts = nullptr;
    // two-step kernels: rows holding a source or receiver per column block,
    // the per-cell event map {ring+1, rx+1, src+1, 0}, the ring-cell tables
    uint8_t *evrow = nullptr;
    double *hilb = nullptr;          // hilbert_table(T) on the device (envelope misfit)
    int hilb_n = 0;
    unsigned *ringtab = nullptr;
    int *pmap = nullptr;
    // time steps per forward / adjoint launch: 1, 2, or 0 = by launch size
    // (two-step launches from pair_groups_fwd / _adj work-groups per CU)
    int fwd_steps = 0, adj_steps = 0;
    int pair_nw = 0;                 // two-step block width requested (0: fewest threads)
    int pair_nw_built = 256;         // the width the two-step tables were built for
    bool points_dirty = true;
    int th_fwd = 32, th_adj = 16, max_batch = 0;
    int pers_fwd = 0, pers_adj = 0;   // persistent-work-group launches (0 = one work-group per tile)
    // cache blocking of the step loops: runs of tb_steps time steps on groups
    // of tb_shots shots (0 = all), so that a group's state stays in the
    // Infinity Cache between consecutive steps; [0] forward, [1] adjoint
    int tb_steps[2] = {1, 1}, tb_shots[2] = {0, 0};
    int n_cu = 0;
    // hipGraph replay of the step loops: slot 0 forward, 1 forward with ring save,
    // 2 adjoint; a graph is reused while its key (shots, tiles, launch style and
    // the epoch of the device buffers it captured) is unchanged
    bool use_graph = true;
    unsigned long long epoch = 1;
    hipGraphExec_t gexec[3] = {};
    unsigned long long gkey[3] = {};
    // loops whose two-step phase splits over stream groups: per slot, the
    // pairs of each group captured on that group's own stream
    hipGraphExec_t gsplit[3][4] = {};
    unsigned long long gskey[3] = {};
    bool gsvalid[3] = {};
    uint8_t *tf_fwd = nullptr, *tf_adj = nullptr;
    int tf_fwd_th = 0, tf_adj_th = 0;
    std::vector<float> obs_h;
    int obs_shots = 0;
    unsigned long long obs_ver = 1;   // bumped by every otfwi_set_observed
    Batch b;
    double *gacc = nullptr;          // FP64 sum of per-shot gradients across batches
    // source-side illumination of the last gradient call: sum over its shots of
    // P(t)^2 at every ill_stride-th step, times ill_stride (0 = not collected)
    int ill_stride = 0;
    float *ill = nullptr;
    bool ill_valid = false;
    // exact discrete adjoint: forward snapshots, the A fields of one segment,
    // the receiver map; FP64 c^2 when given
    void *xsnap = nullptr, *xseg = nullptr;
    size_t xsnap_bytes = 0, xseg_bytes = 0;
    int *rxmap = nullptr;
    int *dflag = nullptr;            // device validity flag of otfwi_set_cquad_device
    std::vector<double> cq64_h;
    float *gout = nullptr;           // FP32 mean gradient on device
    // state of a pending two-stage gradient
    int pend_n = 0;
    // HIP events around the time loops of the last call (forward, misfit,
    // adjoint), accumulated over its batches
    hipEvent_t ev[4] = {};
    // two-step loops can advance two halves of the shots on two streams (the
    // second half's launches fill the tails of the first's); fork / join events
    hipStream_t xs[3] = {};          // extra streams of a split two-step phase
    hipEvent_t evfj[4] = {};         // fork, and the join of each extra stream
    int pair_streams = 0;            // 1, 2, 4, or 0 = by launch size
    double t_fwd_ms = 0, t_mis_ms = 0, t_adj_ms = 0;
    long long n_fwd_launch = 0, n_adj_launch = 0;
};

namespace {

size_t esize(const otfwi_model *m) { return m->prec == 8 ? 8 : 4; }

int ceil_div(int a, int b) { return (a + b - 1) / b; }

void free_batch(Batch &b)
{
    // both slots of a field are one allocation (slot 1 follows slot 0)
    for (void **f : {b.P, b.Ax, b.Az, b.Px, b.LP, b.LAx, b.LAz, b.LPx}) {
        dfree(f[0]);
        f[1] = nullptr;
    }
    dfree(b.grad); dfree(b.ring); dfree(b.rec); dfree(b.adj); dfree(b.jcum);
    dfree(b.lap[0]);
    b.lap[1] = nullptr;
    b.lap_cap = 0;
    dfree(b.mis_tr); dfree(b.mis_shot); dfree(b.scratch); dfree(b.shot_src); dfree(b.obs);
    b.cap = 0;
    b.ring_elems = 0;
    b.obs_nb = 0;
}

int ensure_batch(otfwi_model *m, int n, bool need_adj, bool need_ring = true)
{
    // every run reuses the batch buffers: a pending two-stage gradient
    // (otfwi_forward_for_grad) is no longer valid after another run
    m->pend_n = 0;
    Batch &b = m->b;
    const size_t es = esize(m);
    const size_t plane = (size_t)m->zp * m->ld;
    const size_t ring_need = need_ring ? (size_t)n * m->T * std::max(m->n_ring, 1) : 0;
    const bool adj_ok = b.LP[0] != nullptr || !need_adj;
    if (b.cap >= n && adj_ok && (!need_adj || b.ring_elems >= ring_need)) return OTFWI_OK;
    const int cap = std::max(n, b.cap);
    free_batch(b);
    m->epoch++;
    auto fa = [&](void **p, size_t elems) -> int {
        char *q = nullptr;
        RCK(dalloc(&q, elems * es));
        *p = q;
        return OTFWI_OK;
    };
    // the two slots of a field in one allocation, slot 1 right after slot 0
    // (the two-step kernels address both through one buffer resource)
    auto fa2 = [&](void **f, size_t elems) -> int {
        RCK(fa(&f[0], 2 * elems));
        f[1] = (char *)f[0] + elems * es;
        return OTFWI_OK;
    };
    for (void **f : {b.P, b.Ax, b.Az, b.Px}) RCK(fa2(f, plane * cap));
    if (need_adj)
        for (void **f : {b.LP, b.LAx, b.LAz, b.LPx}) RCK(fa2(f, plane * cap));
    const size_t ntr = (size_t)cap * std::max(m->n_rx, 1);
    RCK(fa(&b.rec, ntr * m->T));
    if (need_adj) {
        RCK(fa(&b.grad, plane * cap));
        b.ring_elems = need_ring ? (size_t)cap * m->T * std::max(m->n_ring, 1) : 0;
        if (b.ring_elems) RCK(fa(&b.ring, b.ring_elems));
        RCK(fa(&b.adj, ntr * m->T));
        RCK(fa(&b.jcum, ntr * m->T));
        RCK(dalloc(&b.mis_tr, ntr));
        RCK(dalloc(&b.mis_shot, (size_t)cap));
        RCK(dalloc(&b.scratch, ntr * 6 * m->T + ntr));   // room for every misfit kind
        RCK(dalloc(&b.obs, ntr * m->T));
    }
    RCK(dalloc(&b.shot_src, (size_t)cap));
    b.cap = cap;
    return OTFWI_OK;
}

// Tile flags for tile height th: STRIP if the E box touches a damped cell,
// EVENT if it holds a source, receiver or ring point.
int build_flags(otfwi_model *m, int th, uint8_t **dst, int *dst_th)
{
    if (*dst && *dst_th == th && !m->points_dirty) return OTFWI_OK;
    const int ntx = ceil_div(m->X, TW), ntz = ceil_div(m->Z, th);
    std::vector<uint8_t> f((size_t)ntx * ntz, 0);
    std::vector<char> ev((size_t)m->X * m->Z, 0);
    auto mark = [&](const std::vector<int> &xs, const std::vector<int> &zs) {
        for (size_t i = 0; i < xs.size(); i++) ev[(size_t)zs[i] * m->X + xs[i]] = 1;
    };
    mark(m->sx, m->sz);
    mark(m->rxx, m->rxz);
    mark(m->vx, m->vz);
    for (int tz = 0; tz < ntz; tz++)
        for (int tx = 0; tx < ntx; tx++) {
            uint8_t fl = 0;
            const int xa = std::max(tx * TW - 4, BOX), xb = std::min(tx * TW + TW + 4, m->X - BOX);
            const int za = std::max(tz * th - 4, BOX), zb = std::min(tz * th + th + 4, m->Z - BOX);
            for (int z = za; z < zb; z++)
                for (int x = xa; x < xb; x++) {
                    if (m->kx_h[x] != 1.0f || m->kz_h[z] != 1.0f) fl |= TF_STRIP;
                    if (ev[(size_t)z * m->X + x]) fl |= TF_EVENT;
                }
            if (tx * TW - 4 < BOX || tx * TW + TW + 4 > m->X - BOX || tz * th - 4 < BOX ||
                tz * th + th + 4 > m->Z - BOX)
                fl |= TF_EDGE;
            f[(size_t)tz * ntx + tx] = fl;
        }
    dfree(*dst);
    RCK(dalloc(dst, f.size()));
    HIPCK(hipMemcpy(*dst, f.data(), f.size(), hipMemcpyHostToDevice));
    *dst_th = th;
    m->epoch++;
    return OTFWI_OK;
}

int check_points(const otfwi_model *m, int n, const int *xs, const int *zs, const char *what)
{
    if (n < 0 || (n > 0 && (!xs || !zs))) return fail(OTFWI_E_ARG, "%s: bad list", what);
    std::vector<char> seen((size_t)m->X * m->Z, 0);
    for (int i = 0; i < n; i++) {
        const int x = xs[i], z = zs[i];
        if (x < BOX || z < BOX || x >= m->X - BOX || z >= m->Z - BOX)
            return fail(OTFWI_E_ARG,
                        "%s %d at (x=%d, z=%d) lies outside the updated interior "
                        "4 <= x < %d, 4 <= z < %d",
                        what, i, x, z, m->X - BOX, m->Z - BOX);
        char &s = seen[(size_t)z * m->X + x];
        if (s) return fail(OTFWI_E_ARG, "%s %d at (x=%d, z=%d) repeats a cell", what, i, x, z);
        s = 1;
    }
    return OTFWI_OK;
}

// Event map and point table from the source, receiver and ring lists.
int build_points(otfwi_model *m)
{
    if (!m->points_dirty && m->evt) return OTFWI_OK;
    std::vector<int> map((size_t)m->zp * m->ld, -1);
    std::vector<PointEvt> pts;
    auto get = [&](int x, int z) -> PointEvt & {
        int &e = map[(size_t)(z + ZOFF) * m->ld + x + XOFF];
        if (e < 0) {
            e = (int)pts.size();
            pts.push_back(PointEvt{-1, -1, -1, 0});
        }
        return pts[e];
    };
    for (int i = 0; i < m->n_src; i++) get(m->sx[i], m->sz[i]).src = i;
    for (int i = 0; i < m->n_rx; i++) get(m->rxx[i], m->rxz[i]).rx = i;
    // ring slots in raster order (z, then x): the two-step reconstruction
    // kernel derives a cell's slot from lane masks and per-word counts
    std::vector<int> ord(m->n_ring);
    for (int i = 0; i < m->n_ring; i++) ord[i] = i;
    std::sort(ord.begin(), ord.end(), [&](int u, int v) {
        return m->vz[u] != m->vz[v] ? m->vz[u] < m->vz[v] : m->vx[u] < m->vx[v];
    });
    for (int r = 0; r < m->n_ring; r++) get(m->vx[ord[r]], m->vz[ord[r]]).ring = r;
    // two-step kernels: rows of each nw-column block (owned columns plus
    // halo) holding a source or receiver, and the per-cell event map
    const int nw = m->pair_nw ? m->pair_nw : f2_width(m->X);
    const int own = f2_own(nw), nwords = nw / 32, rw = f2_ring_words(nw);
    const int nbx = ceil_div(m->X, own);
    std::vector<uint8_t> evrow((size_t)nbx * m->Z, 0);
    std::vector<int> pmap(map.size() * 4, 0);
    for (size_t q = 0; q < map.size(); q++) {
        if (map[q] < 0) continue;
        const PointEvt &pe = pts[map[q]];
        pmap[4 * q] = pe.ring + 1;
        pmap[4 * q + 1] = pe.rx + 1;
        pmap[4 * q + 2] = pe.src + 1;
        if (pe.rx < 0 && pe.src < 0) continue;
        const int z = (int)(q / m->ld) - ZOFF, x = (int)(q % m->ld) - XOFF;
        for (int bx = 0; bx < nbx; bx++) {
            const int c = x - (bx * own - F2_HALO);
            if (c >= 0 && c < nw) evrow[(size_t)bx * m->Z + z] = 1;
        }
    }
    // ring cells per column block and row: {slot of the block's first ring
    // cell, 16-bit count before each 32-lane word, lane masks}
    std::vector<std::vector<int>> ring_rows(m->Z);
    for (int r = 0; r < m->n_ring; r++) ring_rows[m->vz[ord[r]]].push_back(m->vx[ord[r]]);
    std::vector<unsigned> ringtab((size_t)nbx * m->Z * rw, 0u);
    int before = 0;
    for (int z = 0; z < m->Z; z++) {
        const std::vector<int> &row = ring_rows[z];   // ascending x
        for (int bx = 0; bx < nbx; bx++) {
            unsigned *e = &ringtab[((size_t)bx * m->Z + z) * rw];
            const int c0 = bx * own - F2_HALO;
            const int first = (int)(std::lower_bound(row.begin(), row.end(), c0) - row.begin());
            e[0] = (unsigned)(before + first);
            for (int w = 0; w < nwords; w++) {
                const int col0 = c0 + 32 * w;
                const int lo = (int)(std::lower_bound(row.begin(), row.end(), col0) - row.begin());
                unsigned mk = 0;
                for (int j = lo; j < (int)row.size() && row[j] < col0 + 32; j++)
                    mk |= 1u << (row[j] - col0);
                e[1 + w / 2] |= (unsigned)(lo - first) << (16 * (w % 2));   // < nw
                e[1 + nw / 64 + w] = mk;
            }
        }
        before += (int)row.size();
    }
    m->pair_nw_built = nw;
    dfree(m->evrow);
    dfree(m->pmap);
    dfree(m->ringtab);
    RCK(dalloc(&m->ringtab, ringtab.size()));
    HIPCK(hipMemcpy(m->ringtab, ringtab.data(), ringtab.size() * sizeof(unsigned),
                    hipMemcpyHostToDevice));
    RCK(dalloc(&m->evrow, evrow.size()));
    HIPCK(hipMemcpy(m->evrow, evrow.data(), evrow.size(), hipMemcpyHostToDevice));
    RCK(dalloc(&m->pmap, pmap.size()));
    HIPCK(hipMemcpy(m->pmap, pmap.data(), pmap.size() * sizeof(int), hipMemcpyHostToDevice));
    dfree(m->evt);
    dfree(m->pts);
    RCK(dalloc(&m->evt, map.size()));
    HIPCK(hipMemcpy(m->evt, map.data(), map.size() * sizeof(int), hipMemcpyHostToDevice));
    if (!pts.empty()) {
        RCK(dalloc(&m->pts, pts.size()));
        HIPCK(hipMemcpy(m->pts, pts.data(), pts.size() * sizeof(PointEvt), hipMemcpyHostToDevice));
    }
    m->points_dirty = false;
    dfree(m->tf_fwd);
    dfree(m->tf_adj);
    m->epoch++;
    return OTFWI_OK;
}

template <typename R>
int upload_model_arrays(otfwi_model *m)
{
    // cm = (c^2 dt)/dx, rounded as the reference forms it (regression_PML_single.cu:512).
    const size_t n = (size_t)m->zp * m->ld;
    std::vector<float> cq((size_t)m->Z * m->X);
    HIPCK(hipMemcpy(cq.data(), m->cq32, cq.size() * sizeof(float), hipMemcpyDeviceToHost));
    const bool c64 = m->cq64_h.size() == cq.size();
    std::vector<R> cm(n, R(0));
    for (int z = 0; z < m->Z; z++)
        for (int x = 0; x < m->X; x++) {
            const size_t q = (size_t)z * m->X + x;
            const R c = c64 ? (R)m->cq64_h[q] : (R)cq[q];
            cm[(size_t)(z + ZOFF) * m->ld + x + XOFF] = (c * (R)m->dt) / (R)m->dx;
        }
    if (!m->cm) RCK(dalloc((char **)&m->cm, n * sizeof(R)));
    HIPCK(hipMemcpy(m->cm, cm.data(), n * sizeof(R), hipMemcpyHostToDevice));
    return OTFWI_OK;
}

int check_cq(const float *cq, size_t n)
{
    for (size_t i = 0; i < n; i++)
        if (!std::isfinite(cq[i]) || cq[i] < 0.0f)
            return fail(OTFWI_E_NUMERIC, "c^2 at flat index %zu is %g (must be finite and >= 0)",
                        i, (double)cq[i]);
    return OTFWI_OK;
}

int set_cq_host(otfwi_model *m, const float *cq)
{
    if (!cq) return fail(OTFWI_E_ARG, "cquad is NULL");
    const size_t n = (size_t)m->X * m->Z;
    RCK(check_cq(cq, n));
    if (!m->cq32) RCK(dalloc(&m->cq32, n));
    HIPCK(hipMemcpy(m->cq32, cq, n * sizeof(float), hipMemcpyHostToDevice));
    m->have_cq = true;
    m->cq64_h.clear();
    return m->prec == 8 ? upload_model_arrays<double>(m) : upload_model_arrays<float>(m);
}

int set_cq64_host(otfwi_model *m, const double *cq)
{
    if (!cq) return fail(OTFWI_E_ARG, "cquad is NULL");
    const size_t n = (size_t)m->X * m->Z;
    std::vector<float> c32(n);
    for (size_t i = 0; i < n; i++) {
        if (!std::isfinite(cq[i]) || cq[i] < 0.0)
            return fail(OTFWI_E_NUMERIC, "c^2 at flat index %zu is %g (must be finite and >= 0)",
                        i, cq[i]);
        c32[i] = (float)cq[i];
    }
    if (!m->cq32) RCK(dalloc(&m->cq32, n));
    HIPCK(hipMemcpy(m->cq32, c32.data(), n * sizeof(float), hipMemcpyHostToDevice));
    m->have_cq = true;
    m->cq64_h.assign(cq, cq + n);
    return m->prec == 8 ? upload_model_arrays<double>(m) : upload_model_arrays<float>(m);
}

template <typename R>
int upload_sources(otfwi_model *m)
{
    // running sum of each wavelet in the state precision, sequential as in
    // somaFonteIntegral (regression_PML_single.cu:286-291): icum[n][t] =
    // sum_{tau=1..t} s[n][tau]; s[n][0] is never injected.
    std::vector<R> ic((size_t)std::max(m->n_src, 1) * m->T, R(0));
    for (int n = 0; n < m->n_src; n++) {
        R acc = R(0);
        for (int t = 1; t < m->T; t++) {
            acc = acc + (R)m->src_h[(size_t)n * m->T + t];
            ic[(size_t)n * m->T + t] = acc;
        }
    }
    dfree(m->icum);
    m->epoch++;
    RCK(dalloc((char **)&m->icum, ic.size() * sizeof(R)));
    HIPCK(hipMemcpy(m->icum, ic.data(), ic.size() * sizeof(R), hipMemcpyHostToDevice));
    return OTFWI_OK;
}

// k = 1 - d*dt in the state precision from the one-cell-shifted profiles
// d_x[z][x+1] and d_z[z+1][x] (regression_PML_single.cu:505-506, 512-513).
template <typename R>
int upload_damping(otfwi_model *m)
{
    // padded: kx[x + XOFF] over ld entries, kz[z + ZOFF] over zp entries, 1 off the grid
    std::vector<R> kx(m->ld, R(1)), kz(m->zp, R(1));
    m->kx_h.assign(m->X, 1.0f);
    m->kz_h.assign(m->Z, 1.0f);
    for (int x = 0; x < m->X; x++) {
        kx[x + XOFF] = R(1) - (R)m->dxs_h[x] * (R)m->dt;
        m->kx_h[x] = kx[x + XOFF] == R(1) ? 1.0f : 0.0f;
    }
    for (int z = 0; z < m->Z; z++) {
        kz[z + ZOFF] = R(1) - (R)m->dzs_h[z] * (R)m->dt;
        m->kz_h[z] = kz[z + ZOFF] == R(1) ? 1.0f : 0.0f;
    }
    dfree(m->kx);
    dfree(m->kz);
    m->epoch++;
    RCK(dalloc((char **)&m->kx, kx.size() * sizeof(R)));
    RCK(dalloc((char **)&m->kz, kz.size() * sizeof(R)));
    HIPCK(hipMemcpy(m->kx, kx.data(), kx.size() * sizeof(R), hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(m->kz, kz.data(), kz.size() * sizeof(R), hipMemcpyHostToDevice));
    dfree(m->tf_fwd);
    dfree(m->tf_adj);
    return OTFWI_OK;
}

int ready(otfwi_model *m, bool need_ring)
{
    if (!m->have_cq) return fail(OTFWI_E_STATE, "c^2 not set");
    if (!m->have_damp) return fail(OTFWI_E_STATE, "damping not set");
    if (need_ring && m->n_ring == 0) return fail(OTFWI_E_STATE, "boundary ring not set");
    RCK(build_points(m));
    RCK(build_flags(m, m->th_fwd, &m->tf_fwd, &m->tf_fwd_th));
    RCK(build_flags(m, m->th_adj, &m->tf_adj, &m->tf_adj_th));
    return OTFWI_OK;
}

// Work-groups of a persistent launch: every CU filled to the kernel's
// occupancy, rounded down to a multiple of the 8 XCDs (0 = per-tile launch).
// Two steps per launch on this batch?  Explicit 1 / 2, or by size: a
// two-step work-group marches 285 rows in sequence, so a launch needs enough
// of them to fill the CUs (measured on MI355X, 64-step runs: at 64-320
// work-groups one step per launch is 1.05-8x faster; at 640 and more two
// steps are 1.2-1.6x faster; the adjoint's crossover is lower than the
// forward's).
bool use_pairs(const otfwi_model *m, int n, bool adj)
{
    const int want = adj ? m->adj_steps : m->fwd_steps;
    if (want != 0) return want == 2;
    if (m->prec != 4) return false;   // the FP64 two-step kernels run at 1 wave per SIMD
    const long long wg = pair2_groups(m->X, m->Z, m->pair_nw_built, n);
    const long long cu = std::max(m->n_cu, 1);
    return adj ? 2 * wg >= 3 * cu : wg >= 2 * cu;
}

// Streams the pair phase of a loop over n shots is split over (1 or 2):
// two when each half still launches 2 work-groups per CU (measured at
// 64 x 1024^2: two concurrent 32-shot halves take 0.92x the time of one
// 64-shot launch sequence, their launches filling each other's tails).
// A split loop is never captured as one graph across its streams (a process
// ended with a host segmentation fault inside the first forward call of a
// model whose loop had been captured across 2 or 4 streams, with no cause
// found in this code): with graphs on, each shot group's pairs are captured
// on that group's own stream by itself (run_split), and the fork and join
// stay outside the graphs.
int pair_split(const otfwi_model *m, int n)
{
    if (n < 2 || !m->xs[0]) return 1;
    if (m->pair_streams != 0) return m->pair_streams;
    if (m->prec != 4) return 1;
    const long long wg = pair2_groups(m->X, m->Z, m->pair_nw_built, n / 2);
    return wg >= 2LL * std::max(m->n_cu, 1) ? 2 : 1;
}

// Pair arguments of shots [g0, g0 + ng) of the batch (planes, traces, ring,
// adjoint sources, gradient and Laplacian planes offset to the group).
template <typename R>
Pair2Args<R> pair_group(const otfwi_model *m, Pair2Args<R> q, int g0, int ng)
{
    const size_t so = (size_t)g0 * (size_t)q.sstride;
    const size_t to = (size_t)g0 * m->n_rx * m->T, ro = (size_t)g0 * m->T * m->n_ring;
    q.nshots = ng;
    q.P0 += so; q.Ax0 += so; q.Az0 += so;
    q.P1 += so; q.Ax1 += so; q.Az1 += so;
    if (q.Px0) { q.Px0 += so; q.Px1 += so; }
    if (q.shot_src) q.shot_src += g0;
    if (q.rec) q.rec += to;
    if (q.ring) q.ring += ro;
    if (q.ring_in) q.ring_in += ro;
    if (q.jcum) q.jcum += to;
    if (q.grad) q.grad += so;
    if (q.lap0) { q.lap0 += so; q.lap1 += so; }
    return q;
}

// Launches of one pair on the batch's shots, split into ns groups of
// consecutive shots, group k on the model's stream (k = 0) or extra stream
// k - 1 (the caller forks and joins).
template <typename R>
hipError_t launch_pair_split(const otfwi_model *m, const Pair2Args<R> &q, int n, int ns, int mode)
{
    if (ns <= 1) return launch_pair2<R>(q, mode, m->stream);
    for (int k = 0; k < ns; k++) {
        const int g0 = (int)((long long)n * k / ns), g1 = (int)((long long)n * (k + 1) / ns);
        if (g1 <= g0) continue;
        const hipError_t e =
            launch_pair2<R>(pair_group<R>(m, q, g0, g1 - g0), mode, k ? m->xs[k - 1] : m->stream);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Fork: the ns - 1 extra streams start after what is on the model's stream;
// join: the model's stream continues after what is on each of them (all
// recorded into a captured graph).
int fork_streams(otfwi_model *m, int ns)
{
    HIPCK(hipEventRecord(m->evfj[0], m->stream));
    for (int k = 1; k < ns; k++) HIPCK(hipStreamWaitEvent(m->xs[k - 1], m->evfj[0], 0));
    return OTFWI_OK;
}
int join_streams(otfwi_model *m, int ns)
{
    for (int k = 1; k < ns; k++) {
        HIPCK(hipEventRecord(m->evfj[k], m->xs[k - 1]));
        HIPCK(hipStreamWaitEvent(m->stream, m->evfj[k], 0));
    }
    return OTFWI_OK;
}

int persistent_wgs(const otfwi_model *m, bool adj)
{
    if (!(adj ? m->pers_adj : m->pers_fwd) || m->prec == 8) return 0;
    const int occ = step_occupancy(adj, m->prec, adj ? m->th_adj : m->th_fwd, true);
    const int n = std::max(8, (occ * m->n_cu) / 8 * 8);
    return n;
}

// Runs body() (which enqueues kernel launches only) directly, or captures it
// once into a hipGraph for slot `slot` and replays that.  A body that uses the
// extra streams (split) always runs directly.
template <typename F>
int run_loop(otfwi_model *m, int slot, unsigned long long key, bool split, F body)
{
    if (!m->use_graph || split) return body();
    if (!m->gexec[slot] || m->gkey[slot] != key) {
        // the previous replay of this slot may still be running
        if (m->gexec[slot]) {
            HIPCK(hipStreamSynchronize(m->stream));
            (void)hipGraphExecDestroy(m->gexec[slot]);
        }
        m->gexec[slot] = nullptr;
        hipGraph_t graph = nullptr;
        HIPCK(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
        const int rc = body();
        const hipError_t e = hipStreamEndCapture(m->stream, &graph);
        if (rc != OTFWI_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        HIPCK(e);
        const hipError_t ei = hipGraphInstantiate(&m->gexec[slot], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIPCK(ei);
        m->gkey[slot] = key;
    }
    HIPCK(hipGraphLaunch(m->gexec[slot], m->stream));
    return OTFWI_OK;
}

// Shot range [g0, g1) of group k of ns (consecutive shots, as launch_pair_split
// splits them).
inline void group_range(int n, int k, int ns, int &g0, int &g1)
{
    g0 = (int)((long long)n * k / ns);
    g1 = (int)((long long)n * (k + 1) / ns);
}

// A loop whose two-step phase splits over ns stream groups, with graphs on:
// lead() and trail() enqueue the single steps before and after the pairs on
// the model's stream, directly; group(k, s) enqueues the pairs of shot group
// k on stream s.  Each group's pairs are captured once, by themselves, on the
// stream they run on (a one-stream capture, as every other captured loop) and
// replayed there between a fork from and a join into the model's stream; the
// fork and join events are recorded outside the graphs.
template <typename FL, typename FG, typename FT>
int run_split(otfwi_model *m, int slot, unsigned long long key, int n, int ns, FL lead, FG group,
              FT trail)
{
    RCK(lead());
    if (!m->gsvalid[slot] || m->gskey[slot] != key) {
        // earlier replays of this slot may still run on any of the streams
        HIPCK(hipStreamSynchronize(m->stream));
        for (int k = 0; k < 3; k++)
            if (m->xs[k]) HIPCK(hipStreamSynchronize(m->xs[k]));
        for (int k = 0; k < 4; k++) {
            if (m->gsplit[slot][k]) (void)hipGraphExecDestroy(m->gsplit[slot][k]);
            m->gsplit[slot][k] = nullptr;
        }
        m->gsvalid[slot] = false;
        for (int k = 0; k < ns; k++) {
            int g0, g1;
            group_range(n, k, ns, g0, g1);
            if (g1 <= g0) continue;
            hipStream_t s = k ? m->xs[k - 1] : m->stream;
            hipGraph_t graph = nullptr;
            HIPCK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            const int rc = group(k, s);
            const hipError_t e = hipStreamEndCapture(s, &graph);
            if (rc != OTFWI_OK) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            HIPCK(e);
            const hipError_t ei = hipGraphInstantiate(&m->gsplit[slot][k], graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            HIPCK(ei);
        }
        m->gskey[slot] = key;
        m->gsvalid[slot] = true;
    }
    RCK(fork_streams(m, ns));
    for (int k = 0; k < ns; k++)
        if (m->gsplit[slot][k])
            HIPCK(hipGraphLaunch(m->gsplit[slot][k], k ? m->xs[k - 1] : m->stream));
    RCK(join_streams(m, ns));
    return trail();
}

unsigned long long loop_key(const otfwi_model *m, int n, int th, int pw, int t0, int t1)
{
    unsigned long long k = m->epoch;
    k = k * 1000003ULL + (unsigned long long)n;
    k = k * 1000003ULL + (unsigned long long)th;
    k = k * 1000003ULL + (unsigned long long)pw;
    k = k * 1000003ULL + (unsigned long long)t0;
    k = k * 1000003ULL + (unsigned long long)t1;
    return k;
}

// Part of a time loop to enqueue: the whole loop, or (two-step loops only,
// for the fused multi-model gradient) the single steps before the two-step
// phase / after it; the two-step phase itself is then launched by the caller.
enum LoopPart { LOOP_ALL = 0, LOOP_LEAD = 1, LOOP_TRAIL = 2 };
template <typename R>
int fwd_loop(otfwi_model *m, int n, bool save_ring, int t0, int t1, bool illum = false,
             LoopPart part = LOOP_ALL);
template <typename R>
int adj_loop(otfwi_model *m, int n, int t0, int t1, LoopPart part = LOOP_ALL);

template <typename R>
int run_forward(otfwi_model *m, int n, bool save_ring, bool illum = false, LoopPart part = LOOP_ALL)
{
    if (part == LOOP_TRAIL) return fwd_loop<R>(m, n, save_ring, 1, m->T, illum, part);
    Batch &b = m->b;
    const size_t es = sizeof(R);
    const size_t plane = (size_t)m->zp * m->ld;
    hipStream_t st = m->stream;
    for (int i = 0; i < 2; i++) {
        HIPCK(hipMemsetAsync(b.P[i], 0, plane * n * es, st));
        HIPCK(hipMemsetAsync(b.Ax[i], 0, plane * n * es, st));
        HIPCK(hipMemsetAsync(b.Az[i], 0, plane * n * es, st));
        HIPCK(hipMemsetAsync(b.Px[i], 0, plane * n * es, st));
    }
    HIPCK(hipMemsetAsync(b.rec, 0, (size_t)n * m->n_rx * m->T * es, st));
    if (save_ring) {
        // sample 0 of the ring is never written by the steps (reference memsets it)
        for (int s = 0; s < n; s++)
            HIPCK(hipMemsetAsync((char *)b.ring + (size_t)s * m->T * m->n_ring * es, 0,
                                 (size_t)m->n_ring * es, st));
    }
    return fwd_loop<R>(m, n, save_ring, 1, m->T, illum, part);
}

// Launches of forward steps t0 .. t1-1 on the current batch (no memsets).
// Step-invariant arguments of the forward step on the current batch of n shots.
template <typename R>
FwdArgs<R> fwd_args(otfwi_model *m, int n)
{
    Batch &b = m->b;
    FwdArgs<R> a{};
    a.X = m->X; a.Z = m->Z; a.ld = m->ld;
    a.ntx = ceil_div(m->X, TW); a.ntz = ceil_div(m->Z, m->th_fwd);
    a.nshots = n; a.T = m->T;
    a.sstride = (long long)m->zp * m->ld;
    a.cm = (const R *)m->cm;
    a.cA = (R)m->dt / (R)m->dx;
    a.kx = (const R *)m->kx; a.kz = (const R *)m->kz;
    a.tflags = m->tf_fwd;
    a.evt = m->evt; a.pts = m->pts;
    a.shot_src = b.shot_src;
    a.icum = (const R *)m->icum;
    a.rec = (R *)b.rec; a.n_rx = m->n_rx;
    a.ring = (R *)b.ring;
    a.n_ring = m->n_ring;
    return a;
}

// Step-invariant arguments of the two-step kernels on the current batch of n shots.
template <typename R>
Pair2Args<R> pair_args(otfwi_model *m, int n)
{
    Pair2Args<R> q{};
    q.X = m->X; q.Z = m->Z; q.ld = m->ld; q.nshots = n; q.T = m->T;
    q.n_rx = m->n_rx; q.n_ring = m->n_ring;
    q.sstride = (long long)m->zp * m->ld;
    q.cm = (const R *)m->cm; q.kx = (const R *)m->kx; q.kz = (const R *)m->kz;
    q.cA = (R)m->dt / (R)m->dx;
    q.slot_elems = (long long)m->b.cap * q.sstride;
    q.nw = m->pair_nw_built;
    return q;
}

// mode 0 (forward field) from the one-step forward arguments
template <typename R>
Pair2Args<R> pair_fwd_args(otfwi_model *m, int n, const FwdArgs<R> &a)
{
    Pair2Args<R> q = pair_args<R>(m, n);
    q.shot_src = a.shot_src; q.icum = a.icum; q.rec = a.rec; q.ring = a.ring;
    q.mask = m->ringtab; q.evrow = m->evrow; q.pmap = m->pmap;
    return q;
}

// mode 2 (reconstruction) and mode 1 (lambda); slot pointers are set per launch
template <typename R>
Pair2Args<R> pair_rec_args(otfwi_model *m, int n)
{
    Pair2Args<R> u = pair_args<R>(m, n);
    Batch &b = m->b;
    u.lap0 = (R *)b.lap[0]; u.lap1 = (R *)b.lap[1];
    u.ring_in = (const R *)b.ring; u.mask = m->ringtab;
    return u;
}
template <typename R>
Pair2Args<R> pair_lam_args(otfwi_model *m, int n)
{
    Pair2Args<R> l = pair_args<R>(m, n);
    Batch &b = m->b;
    l.lap0 = (R *)b.lap[0]; l.lap1 = (R *)b.lap[1];
    l.jcum = (const R *)b.jcum; l.grad = (R *)b.grad; l.evrow = m->evrow; l.pmap = m->pmap;
    return l;
}

// the two-step adjoint's lap planes (two per shot of the batch)
template <typename R>
int ensure_lap(otfwi_model *m)
{
    Batch &b = m->b;
    if (b.lap[0] && b.lap_cap >= b.cap) return OTFWI_OK;
    dfree(b.lap[0]);
    b.lap[1] = nullptr;
    b.lap_cap = 0;
    const size_t plane = (size_t)m->zp * m->ld;
    char *q = nullptr;   // both planes in one allocation, like the field slots
    RCK(dalloc(&q, 2 * plane * b.cap * sizeof(R)));
    b.lap[0] = q;
    b.lap[1] = q + plane * b.cap * sizeof(R);
    b.lap_cap = b.cap;
    m->epoch++;
    return OTFWI_OK;
}

// Two-step schedule of forward steps t0 .. t1-1: nlead single steps, then
// npairs (even) two-step launches (each writes the slot it does not read, so
// an even count leaves state t in slot t&1 again), then single steps -- the
// last step is always single, so A(t1-2) stays in slot (t1-2)&1 for the
// adjoint's reconstruction.
inline void fwd_schedule(int t0, int t1, int &nlead, int &npairs)
{
    const int S = t1 - t0;
    npairs = S > 1 ? (S - 1) / 2 : 0;
    npairs -= npairs & 1;
    nlead = (S - 1) - 2 * npairs;
}

// Two-step schedule of adjoint steps t0 .. t1-1: nlead single steps, then
// npairs (even) pairs of launches, so the slots end where one step per launch
// leaves them.
inline void adj_schedule(int t0, int t1, int &nlead, int &npairs)
{
    const int S = t1 - t0;
    npairs = (S / 2) & ~1;
    nlead = S - 2 * npairs;
}

// Slot pointers of forward pair k starting at step t: each pair writes the
// slot it does not read, so after k pairs the state sits in the other slot
// than one step per launch leaves it.
template <typename R>
void fwd_pair_slots(const otfwi_model *m, Pair2Args<R> &q, int k, int t)
{
    const Batch &b = m->b;
    const int in = ((t - 1) & 1) ^ (k & 1), out = in ^ 1;
    q.t = t;
    q.P0 = (const R *)b.P[in]; q.P1 = (R *)b.P[out];
    q.Ax0 = (const R *)b.Ax[in]; q.Ax1 = (R *)b.Ax[out];
    q.Az0 = (const R *)b.Az[in]; q.Az1 = (R *)b.Az[out];
    q.Px0 = (const R *)b.Px[in]; q.Px1 = (R *)b.Px[out];
}

// Slot pointers of adjoint pair k starting at step t, reconstruction (u) and
// lambda (l) launches: each pair writes the slots it does not read.  The
// reconstructed field U(t) lives in P slot (T-1+t)%2 and A_b(t) in slot
// (T-2+t)%2, starting from the forward's P(T-1) and A(T-2)
// (regression_PML_single.cu:787-790).
template <typename R>
void adj_pair_slots(const otfwi_model *m, Pair2Args<R> &u, Pair2Args<R> &l, int k, int t)
{
    const Batch &b = m->b;
    const int T = m->T;
    const int f = k & 1;
    const int li = ((t - 1) & 1) ^ f, ui = ((T - 2 + t) & 1) ^ f, ai = ((T - 1 + t) & 1) ^ f;
    u.t = l.t = t;
    u.P0 = (const R *)b.P[ui]; u.P1 = (R *)b.P[ui ^ 1];
    u.Ax0 = (const R *)b.Ax[ai]; u.Ax1 = (R *)b.Ax[ai ^ 1];
    u.Az0 = (const R *)b.Az[ai]; u.Az1 = (R *)b.Az[ai ^ 1];
    l.P0 = (const R *)b.LP[li]; l.P1 = (R *)b.LP[li ^ 1];
    l.Ax0 = (const R *)b.LAx[li]; l.Ax1 = (R *)b.LAx[li ^ 1];
    l.Az0 = (const R *)b.LAz[li]; l.Az1 = (R *)b.LAz[li ^ 1];
    l.Px0 = (const R *)b.LPx[li]; l.Px1 = (R *)b.LPx[li ^ 1];
}

template <typename R>
int fwd_loop(otfwi_model *m, int n, bool save_ring, int t0, int t1, bool illum, LoopPart part)
{
    Batch &b = m->b;
    const size_t plane = (size_t)m->zp * m->ld;
    hipStream_t st = m->stream;
    FwdArgs<R> a = fwd_args<R>(m, n);
    if (!save_ring) a.ring = nullptr;
    const int pw = persistent_wgs(m, false);
    const int ist = (illum && m->ill) ? m->ill_stride : 0;
    const int K = std::max(1, m->tb_steps[0]);
    const int G = m->tb_shots[0] > 0 ? std::min(m->tb_shots[0], n) : n;
    // two steps per launch, scheduled by fwd_schedule.  With illumination
    // sampled every ist-th step (ist even) the pairs start at an odd step, so
    // every sampled step is a single step or a pair's second step, whose P
    // plane the pair writes; such a loop runs on one stream (the illumination
    // adds every shot into one plane).
    const bool two = use_pairs(m, n, false) && ist % 2 == 0 && K == 1 && G == n && m->evrow &&
                     pair2_supported((long long)plane, (long long)b.cap * (long long)plane, (int)sizeof(R));
    int nlead, npairs;
    fwd_schedule(t0, t1, nlead, npairs);
    if (ist > 0 && ((t0 + nlead) & 1) == 0) {
        // one more leading single step and two pairs fewer: the pair count
        // stays even and the last step single (four trailing single steps)
        if (npairs >= 2) {
            npairs -= 2;
            nlead += 1;
        } else {
            nlead += 2 * npairs;
            npairs = 0;
        }
    }
    auto single = [&](int tt) -> int {
        const int p = (tt - 1) & 1, c = tt & 1;
        a.t = tt;
        a.P0 = (const R *)b.P[p]; a.P1 = (R *)b.P[c];
        a.Ax0 = (const R *)b.Ax[p]; a.Ax1 = (R *)b.Ax[c];
        a.Az0 = (const R *)b.Az[p]; a.Az1 = (R *)b.Az[c];
        a.Px0 = (const R *)b.Px[p]; a.Px1 = (R *)b.Px[c];
        HIPCK(launch_fwd_step<R>(a, m->th_fwd, pw, st));
        if (ist > 0 && tt % ist == 0)
            HIPCK(launch_illum<R>((const R *)b.P[c], n, m->X, m->Z, m->ld, (long long)plane,
                                  (float)ist, m->ill, st));
        return OTFWI_OK;
    };
    auto lead_fn = [&]() -> int {
        for (int t = t0; t < t0 + nlead; t++) RCK(single(t));
        return OTFWI_OK;
    };
    auto trail_fn = [&]() -> int {
        for (int t = t0 + std::max(nlead, 0) + 2 * npairs; t < t1; t++) RCK(single(t));
        return OTFWI_OK;
    };
    auto pair_slots = [&](Pair2Args<R> &q, int k, int t) { fwd_pair_slots<R>(m, q, k, t); };
    // all pairs of shot group g of ns on stream s (captured by run_split)
    auto group_fn = [&](int g, int ns, hipStream_t s) -> int {
        int g0, g1;
        group_range(n, g, ns, g0, g1);
        Pair2Args<R> q = pair_fwd_args<R>(m, n, a);
        int t = t0 + nlead;
        for (int k = 0; k < npairs; k++, t += 2) {
            pair_slots(q, k, t);
            HIPCK(launch_pair2<R>(pair_group<R>(m, q, g0, g1 - g0), 0, s));
        }
        return OTFWI_OK;
    };
    // direct launches: the groups' launches of each pair enqueued together
    auto body2 = [&]() -> int {
        RCK(lead_fn());
        Pair2Args<R> q = pair_fwd_args<R>(m, n, a);
        const int ns = ist > 0 ? 1 : pair_split(m, n);
        if (ns > 1 && npairs > 0) RCK(fork_streams(m, ns));
        int t = t0 + nlead;
        for (int k = 0; k < npairs; k++, t += 2) {
            pair_slots(q, k, t);
            HIPCK(launch_pair_split<R>(m, q, n, ns, 0));
            if (ist > 0 && (t + 1) % ist == 0)
                HIPCK(launch_illum<R>((const R *)q.P1, n, m->X, m->Z, m->ld, (long long)plane,
                                      (float)ist, m->ill, st));
        }
        if (ns > 1 && npairs > 0) RCK(join_streams(m, ns));
        return trail_fn();
    };
    auto body = [&]() -> int {
        if (two) return body2();
        for (int tb = t0; tb < t1; tb += K) {
            const int te = std::min(tb + K, t1);
            for (int g0 = 0; g0 < n; g0 += G) {
                const int ng = std::min(G, n - g0);
                const size_t so = (size_t)g0 * plane;               // plane offset of the group
                const size_t to = (size_t)g0 * m->n_rx * m->T;      // trace offset
                FwdArgs<R> ag = a;
                ag.nshots = ng;
                ag.shot_src = b.shot_src + g0;
                ag.rec = a.rec + to;
                ag.ring = a.ring ? a.ring + (size_t)g0 * m->T * m->n_ring : nullptr;
                for (int t = tb; t < te; t++) {
                    const int p = (t - 1) & 1, c = t & 1;
                    ag.t = t;
                    ag.P0 = (const R *)b.P[p] + so; ag.P1 = (R *)b.P[c] + so;
                    ag.Ax0 = (const R *)b.Ax[p] + so; ag.Ax1 = (R *)b.Ax[c] + so;
                    ag.Az0 = (const R *)b.Az[p] + so; ag.Az1 = (R *)b.Az[c] + so;
                    ag.Px0 = (const R *)b.Px[p] + so; ag.Px1 = (R *)b.Px[c] + so;
                    HIPCK(launch_fwd_step<R>(ag, m->th_fwd, pw, st));
                    if (ist > 0 && t % ist == 0)
                        HIPCK(launch_illum<R>((const R *)b.P[c] + so, ng, m->X, m->Z, m->ld,
                                              (long long)plane, (float)ist, m->ill, st));
                }
            }
        }
        return OTFWI_OK;
    };
    const unsigned long long kb =
        (((unsigned long long)K * 4099ULL + (unsigned long long)G) * 3ULL + (two ? 1ULL : 0ULL)) *
            3ULL + (unsigned long long)pair_split(m, n);
    const unsigned long long key =
        (loop_key(m, n, m->th_fwd, pw, t0, t1) * 31ULL + (unsigned long long)ist) * 1000003ULL + kb;
    if (part != LOOP_ALL) {
        if (!two) return fail(OTFWI_E_STATE, "forward loop runs one step per launch");
        return part == LOOP_LEAD ? lead_fn() : trail_fn();
    }
    const int ns = two && npairs > 0 && ist == 0 ? pair_split(m, n) : 1;
    if (ns > 1 && m->use_graph)
        return run_split(m, save_ring ? 1 : 0, key, n, ns, lead_fn,
                         [&](int g, hipStream_t s) { return group_fn(g, ns, s); }, trail_fn);
    return run_loop(m, save_ring ? 1 : 0, key, ns > 1, body);
}

template <typename R>
int run_adjoint(otfwi_model *m, int n, LoopPart part = LOOP_ALL)
{
    Batch &b = m->b;
    const size_t es = sizeof(R);
    const size_t plane = (size_t)m->zp * m->ld;
    hipStream_t st = m->stream;
    for (int i = 0; i < 2; i++) {
        HIPCK(hipMemsetAsync(b.LP[i], 0, plane * n * es, st));
        HIPCK(hipMemsetAsync(b.LAx[i], 0, plane * n * es, st));
        HIPCK(hipMemsetAsync(b.LAz[i], 0, plane * n * es, st));
        HIPCK(hipMemsetAsync(b.LPx[i], 0, plane * n * es, st));
    }
    HIPCK(hipMemsetAsync(b.grad, 0, plane * n * es, st));
    return adj_loop<R>(m, n, 1, m->T, part);
}

// Launches of adjoint steps t0 .. t1-1 on the current batch (no memsets).
// Step-invariant arguments of the one-step adjoint on the current batch of n shots.
template <typename R>
AdjArgs<R> adj_args(otfwi_model *m, int n)
{
    Batch &b = m->b;
    AdjArgs<R> a{};
    a.X = m->X; a.Z = m->Z; a.ld = m->ld;
    a.ntx = ceil_div(m->X, TW); a.ntz = ceil_div(m->Z, m->th_adj);
    a.nshots = n; a.T = m->T;
    a.sstride = (long long)m->zp * m->ld;
    a.grad = (R *)b.grad;
    a.cm = (const R *)m->cm;
    a.cA = (R)m->dt / (R)m->dx;
    a.kx = (const R *)m->kx; a.kz = (const R *)m->kz;
    a.tflags = m->tf_adj;
    a.evt = m->evt; a.pts = m->pts;
    a.jcum = (const R *)b.jcum; a.n_rx = m->n_rx;
    a.ring = (const R *)b.ring; a.n_ring = m->n_ring;
    return a;
}

template <typename R>
int adj_loop(otfwi_model *m, int n, int t0, int t1, LoopPart part)
{
    Batch &b = m->b;
    const size_t plane = (size_t)m->zp * m->ld;
    hipStream_t st = m->stream;
    AdjArgs<R> a = adj_args<R>(m, n);
    const int T = m->T;
    const int pw = persistent_wgs(m, true);
    const int K = std::max(1, m->tb_steps[1]);
    const int G = m->tb_shots[1] > 0 ? std::min(m->tb_shots[1], n) : n;
    // two steps per launch pair: the reconstruction kernel advances U and
    // writes lap(U) of both steps, then the lambda kernel advances lambda and
    // adds both gradient terms.  Leading single steps, then an even number of
    // pairs (each writes the slots it does not read), so the slots end where
    // one step per launch leaves them.
    const bool two = use_pairs(m, n, true) && K == 1 && G == n && m->evrow && m->ringtab &&
                     pair2_supported((long long)plane, (long long)b.cap * (long long)plane, (int)sizeof(R));
    if (two) RCK(ensure_lap<R>(m));
    auto single = [&](int t) -> int {
        const int lp = (t - 1) & 1, lc = t & 1;
        const int up = (T - 2 + t) & 1, uc = (T - 1 + t) & 1;
        const int ap = (T - 1 + t) & 1, ac = (T - 2 + t) & 1;
        a.t = t;
        a.LP0 = (const R *)b.LP[lp]; a.LP1 = (R *)b.LP[lc];
        a.LAx0 = (const R *)b.LAx[lp]; a.LAx1 = (R *)b.LAx[lc];
        a.LAz0 = (const R *)b.LAz[lp]; a.LAz1 = (R *)b.LAz[lc];
        a.LPx0 = (const R *)b.LPx[lp]; a.LPx1 = (R *)b.LPx[lc];
        a.U0 = (const R *)b.P[up]; a.U1 = (R *)b.P[uc];
        a.UAx0 = (const R *)b.Ax[ap]; a.UAx1 = (R *)b.Ax[ac];
        a.UAz0 = (const R *)b.Az[ap]; a.UAz1 = (R *)b.Az[ac];
        HIPCK(launch_adj_step<R>(a, m->th_adj, pw, st));
        return OTFWI_OK;
    };
    int nlead, npairs;
    adj_schedule(t0, t1, nlead, npairs);
    auto lead_fn = [&]() -> int {
        for (int t = t0; t < t0 + nlead; t++) RCK(single(t));
        return OTFWI_OK;
    };
    auto pair_slots = [&](Pair2Args<R> &u, Pair2Args<R> &l, int k, int t) {
        adj_pair_slots<R>(m, u, l, k, t);
    };
    // all pairs of shot group g of ns on stream s (captured by run_split)
    auto group_fn = [&](int g, int ns, hipStream_t s) -> int {
        int g0, g1;
        group_range(n, g, ns, g0, g1);
        Pair2Args<R> u = pair_rec_args<R>(m, n), l = pair_lam_args<R>(m, n);
        int t = t0 + nlead;
        for (int k = 0; k < npairs; k++, t += 2) {
            pair_slots(u, l, k, t);
            HIPCK(launch_pair2<R>(pair_group<R>(m, u, g0, g1 - g0), 2, s));
            HIPCK(launch_pair2<R>(pair_group<R>(m, l, g0, g1 - g0), 1, s));
        }
        return OTFWI_OK;
    };
    // direct launches: the groups' launches of each pair enqueued together
    auto body2 = [&]() -> int {
        RCK(lead_fn());
        Pair2Args<R> u = pair_rec_args<R>(m, n), l = pair_lam_args<R>(m, n);
        const int ns = pair_split(m, n);
        if (ns > 1 && npairs > 0) RCK(fork_streams(m, ns));
        int t = t0 + nlead;
        for (int k = 0; k < npairs; k++, t += 2) {
            pair_slots(u, l, k, t);
            HIPCK(launch_pair_split<R>(m, u, n, ns, 2));
            HIPCK(launch_pair_split<R>(m, l, n, ns, 1));
        }
        if (ns > 1 && npairs > 0) RCK(join_streams(m, ns));
        return OTFWI_OK;
    };
    auto body = [&]() -> int {
        if (two) return body2();
        for (int tb = t0; tb < t1; tb += K) {
            const int te = std::min(tb + K, t1);
            for (int g0 = 0; g0 < n; g0 += G) {
                const int ng = std::min(G, n - g0);
                const size_t so = (size_t)g0 * plane;
                AdjArgs<R> ag = a;
                ag.nshots = ng;
                ag.grad = a.grad + so;
                ag.jcum = a.jcum + (size_t)g0 * m->n_rx * T;
                ag.ring = a.ring + (size_t)g0 * T * m->n_ring;
                for (int t = tb; t < te; t++) {
                    const int lp = (t - 1) & 1, lc = t & 1;
                    // reconstructed field: U(t) lives in P slot (T-1+t)%2, A_b(t) in slot
                    // (T-2+t)%2, starting from the forward's P(T-1) and A(T-2)
                    // (regression_PML_single.cu:787-790)
                    const int up = (T - 2 + t) & 1, uc = (T - 1 + t) & 1;
                    const int ap = (T - 1 + t) & 1, ac = (T - 2 + t) & 1;
                    ag.t = t;
                    ag.LP0 = (const R *)b.LP[lp] + so; ag.LP1 = (R *)b.LP[lc] + so;
                    ag.LAx0 = (const R *)b.LAx[lp] + so; ag.LAx1 = (R *)b.LAx[lc] + so;
                    ag.LAz0 = (const R *)b.LAz[lp] + so; ag.LAz1 = (R *)b.LAz[lc] + so;
                    ag.LPx0 = (const R *)b.LPx[lp] + so; ag.LPx1 = (R *)b.LPx[lc] + so;
                    ag.U0 = (const R *)b.P[up] + so; ag.U1 = (R *)b.P[uc] + so;
                    ag.UAx0 = (const R *)b.Ax[ap] + so; ag.UAx1 = (R *)b.Ax[ac] + so;
                    ag.UAz0 = (const R *)b.Az[ap] + so; ag.UAz1 = (R *)b.Az[ac] + so;
                    HIPCK(launch_adj_step<R>(ag, m->th_adj, pw, st));
                }
            }
        }
        return OTFWI_OK;
    };
    const unsigned long long kb =
        (((unsigned long long)K * 4099ULL + (unsigned long long)G) * 3ULL + (two ? 1ULL : 0ULL)) *
            3ULL + (unsigned long long)pair_split(m, n);
    const unsigned long long key = loop_key(m, n, m->th_adj, pw, t0, t1) * 1000003ULL + kb;
    if (part != LOOP_ALL) {
        if (!two) return fail(OTFWI_E_STATE, "adjoint loop runs one step per launch");
        return part == LOOP_LEAD ? lead_fn() : OTFWI_OK;
    }
    const int ns = two && npairs > 0 ? pair_split(m, n) : 1;
    if (ns > 1 && m->use_graph)
        return run_split(m, 2, key, n, ns, lead_fn,
                         [&](int g, hipStream_t s) { return group_fn(g, ns, s); },
                         []() { return (int)OTFWI_OK; });
    return run_loop(m, 2, key, ns > 1, body);
}

int upload_shot_src(otfwi_model *m, int n, const int *shot_src)
{
    std::vector<int> ss(n, -1);
    for (int i = 0; i < n; i++) {
        ss[i] = shot_src ? shot_src[i] : -1;
        if (ss[i] < -1 || ss[i] >= m->n_src)
            return fail(OTFWI_E_ARG, "shot %d fires source %d (have %d sources)", i, ss[i], m->n_src);
    }
    HIPCK(hipMemcpyAsync(m->b.shot_src, ss.data(), n * sizeof(int), hipMemcpyHostToDevice,
                         m->stream));
    HIPCK(hipStreamSynchronize(m->stream));
    return OTFWI_OK;
}

template <typename R>
int copy_rec_out(otfwi_model *m, int n, float *rec_out)
{
    const size_t cnt = (size_t)n * m->n_rx * m->T;
    if (sizeof(R) == 4) {
        HIPCK(hipMemcpyAsync(rec_out, m->b.rec, cnt * sizeof(float), hipMemcpyDeviceToHost,
                             m->stream));
        HIPCK(hipStreamSynchronize(m->stream));
    } else {
        std::vector<double> tmp(cnt);
        HIPCK(hipMemcpyAsync(tmp.data(), m->b.rec, cnt * sizeof(double), hipMemcpyDeviceToHost,
                             m->stream));
        HIPCK(hipStreamSynchronize(m->stream));
        for (size_t i = 0; i < cnt; i++) rec_out[i] = (float)tmp[i];
    }
    return OTFWI_OK;
}

int batch_size(otfwi_model *m, int n)
{
    return m->max_batch > 0 ? std::min(n, m->max_batch) : n;
}

template <typename R>
int forward_impl(otfwi_model *m, int n, const int *shot_src, float *rec_out)
{
    RCK(ready(m, false));
    const int B = batch_size(m, n);
    RCK(ensure_batch(m, B, false));
    for (int s0 = 0; s0 < n; s0 += B) {
        const int nb = std::min(B, n - s0);
        RCK(upload_shot_src(m, nb, shot_src ? shot_src + s0 : nullptr));
        RCK(run_forward<R>(m, nb, false));
        if (rec_out) RCK(copy_rec_out<R>(m, nb, rec_out + (size_t)s0 * m->n_rx * m->T));
    }
    HIPCK(hipStreamSynchronize(m->stream));
    return OTFWI_OK;
}

template <typename R>
int upload_obs_batch(otfwi_model *m, int s0, int nb)
{
    if (m->obs_shots < s0 + nb)
        return fail(OTFWI_E_STATE, "observed data holds %d shots, need %d", m->obs_shots, s0 + nb);
    Batch &b = m->b;
    // the traces stay resident while the same shots of the same data are asked for
    if (b.obs_nb == nb && b.obs_s0 == s0 && b.obs_ver == m->obs_ver) return OTFWI_OK;
    const size_t per = (size_t)m->n_rx * m->T;
    HIPCK(hipMemcpyAsync(b.obs, m->obs_h.data() + (size_t)s0 * per, nb * per * sizeof(float),
                         hipMemcpyHostToDevice, m->stream));
    b.obs_s0 = s0;
    b.obs_nb = nb;
    b.obs_ver = m->obs_ver;
    return OTFWI_OK;
}


int check_w2_flags(const otfwi_model *m, const std::vector<int> &fl, int s0)
{
    for (size_t i = 0; i < fl.size(); i++)
        if (fl[i])
            return fail(OTFWI_E_NUMERIC,
                        fl[i] == 2 ? "W2: observed trace of receiver %zu, shot %d is all zero "
                                     "(the W2 normalization is 0/0)"
                                   : "W2: observed trace of receiver %zu, shot %d has a "
                                     "non-finite CDF",
                        i % m->n_rx, s0 + (int)(i / m->n_rx));
    return OTFWI_OK;
}

bool misfit_kind_ok(int k)
{
    return k >= OTFWI_MISFIT_L2 && k <= OTFWI_MISFIT_W2_YANG;
}

// W2 and the four further Wasserstein-type misfits report per-trace flags.
bool w2_family(int k)
{
    return k == OTFWI_MISFIT_W2 || (k >= OTFWI_MISFIT_W2_QUAD && k <= OTFWI_MISFIT_W2_YANG);
}

// The envelope misfit's Hilbert kernel for this model's T, on the device.
int ensure_hilb(otfwi_model *m)
{
    if (m->hilb && m->hilb_n == m->T) return OTFWI_OK;
    std::vector<double> c(m->T);
    hilbert_table(m->T, c.data());
    dfree(m->hilb);
    m->hilb = nullptr;
    m->hilb_n = 0;
    RCK(dalloc(&m->hilb, c.size()));
    HIPCK(hipMemcpy(m->hilb, c.data(), c.size() * sizeof(double), hipMemcpyHostToDevice));
    m->hilb_n = m->T;
    return OTFWI_OK;
}

// One batch of a gradient: forward with ring save, misfit + adjoint source
// (device kind) or host adjoint source already in b.adj (kind < 0), adjoint
// run, and accumulation of the batch's per-shot gradients into m->gacc.
// Misfit of the batch's traces against its observed traces on the device:
// per-trace values, adjoint sources and their running sums, per-shot sums.
template <typename R>
int batch_misfit(otfwi_model *m, int nb, int kind)
{
    Batch &b = m->b;
    const int ntr = nb * m->n_rx;
    int *flags = misfit_flags(kind, b.scratch, ntr, m->T);
    if (w2_family(kind)) HIPCK(hipMemsetAsync(flags, 0, ntr * sizeof(int), m->stream));
    HIPCK((launch_misfit<R, float>(kind, (const R *)b.rec, b.obs, (R *)b.adj, (R *)b.jcum,
                                   b.mis_tr, ntr, m->T, b.scratch, m->stream, m->hilb)));
    HIPCK(launch_reduce_misfit(b.mis_tr, nb, m->n_rx, b.mis_shot, m->stream));
    return OTFWI_OK;
}

template <typename R>
int grad_batch_finish(otfwi_model *m, int s0, int nb, int kind, double *mis_out);

template <typename R>
int grad_batch_device(otfwi_model *m, int s0, int nb, const int *shot_src, int kind,
                      double *mis_out, float *rec_out)
{
    RCK(upload_shot_src(m, nb, shot_src ? shot_src + s0 : nullptr));
    RCK(upload_obs_batch<R>(m, s0, nb));
    HIPCK(hipEventRecord(m->ev[0], m->stream));
    RCK(run_forward<R>(m, nb, true, true));
    HIPCK(hipEventRecord(m->ev[1], m->stream));
    if (rec_out) RCK(copy_rec_out<R>(m, nb, rec_out + (size_t)s0 * m->n_rx * m->T));
    RCK(batch_misfit<R>(m, nb, kind));
    HIPCK(hipEventRecord(m->ev[2], m->stream));
    RCK(run_adjoint<R>(m, nb));
    HIPCK(hipEventRecord(m->ev[3], m->stream));
    return grad_batch_finish<R>(m, s0, nb, kind, mis_out);
}

// After a batch's adjoint: its gradients added to the FP64 sum, per-shot
// misfits and W2 flags read back, phase times taken from ev[0..3].
template <typename R>
int grad_batch_finish(otfwi_model *m, int s0, int nb, int kind, double *mis_out)
{
    Batch &b = m->b;
    const int ntr = nb * m->n_rx;
    int *flags = misfit_flags(kind, b.scratch, ntr, m->T);
    const size_t plane = (size_t)m->zp * m->ld;
    HIPCK(launch_accum_grad<R>((const R *)b.grad, nb, m->X, m->Z, m->ld, (long long)plane, m->gacc,
                               m->stream));
    std::vector<double> hmis(nb);
    HIPCK(hipMemcpyAsync(hmis.data(), b.mis_shot, nb * sizeof(double), hipMemcpyDeviceToHost,
                         m->stream));
    std::vector<int> fl(w2_family(kind) ? ntr : 0);
    if (w2_family(kind))
        HIPCK(hipMemcpyAsync(fl.data(), flags, ntr * sizeof(int), hipMemcpyDeviceToHost, m->stream));
    HIPCK(hipStreamSynchronize(m->stream));
    float ms[3];
    HIPCK(hipEventElapsedTime(&ms[0], m->ev[0], m->ev[1]));
    HIPCK(hipEventElapsedTime(&ms[1], m->ev[1], m->ev[2]));
    HIPCK(hipEventElapsedTime(&ms[2], m->ev[2], m->ev[3]));
    m->t_fwd_ms += ms[0];
    m->t_mis_ms += ms[1];
    m->t_adj_ms += ms[2];
    m->n_fwd_launch += m->T - 1;
    m->n_adj_launch += m->T - 1;
    RCK(check_w2_flags(m, fl, s0));
    if (mis_out)
        for (int i = 0; i < nb; i++) mis_out[s0 + i] = hmis[i];
    return OTFWI_OK;
}

int ensure_ill(otfwi_model *m)
{
    m->ill_valid = false;
    if (m->ill_stride <= 0) return OTFWI_OK;
    const size_t n = (size_t)m->X * m->Z;
    if (!m->ill) {
        RCK(dalloc(&m->ill, n));
        m->epoch++;
    }
    HIPCK(hipMemsetAsync(m->ill, 0, n * sizeof(float), m->stream));
    return OTFWI_OK;
}

int ensure_gacc(otfwi_model *m)
{
    const size_t n = (size_t)m->X * m->Z;
    if (!m->gacc) RCK(dalloc(&m->gacc, n));
    if (!m->gout) RCK(dalloc(&m->gout, n));
    HIPCK(hipMemsetAsync(m->gacc, 0, n * sizeof(double), m->stream));
    return OTFWI_OK;
}

// Checks and device state of a gradient call over n shots (batches of *B).
int grad_prepare(otfwi_model *m, int n, int kind, int *B)
{
    if (!misfit_kind_ok(kind)) return fail(OTFWI_E_ARG, "unknown misfit kind %d", kind);
    if (kind == OTFWI_MISFIT_ENV) RCK(ensure_hilb(m));
    if (n <= 0) return fail(OTFWI_E_ARG, "n_shots must be positive");
    if (m->n_rx <= 0) return fail(OTFWI_E_STATE, "receivers not set");
    if (w2_family(kind) && m->T < 2) return fail(OTFWI_E_ARG, "W2 needs T >= 2");
    RCK(ready(m, true));
    *B = batch_size(m, n);
    RCK(ensure_batch(m, *B, true));
    RCK(ensure_gacc(m));
    RCK(ensure_ill(m));
    m->t_fwd_ms = m->t_mis_ms = m->t_adj_ms = 0;
    m->n_fwd_launch = m->n_adj_launch = 0;
    return OTFWI_OK;
}

// The mean gradient over n shots from the FP64 sum, to grad_out (host or device).
int grad_complete(otfwi_model *m, int n, float *grad_out, bool out_dev)
{
    const int cells = m->X * m->Z;
    HIPCK(launch_finalize_grad(m->gacc, cells, 1.0 / n, m->gout, m->stream));
    if (grad_out) {
        HIPCK(hipMemcpyAsync(grad_out, m->gout, cells * sizeof(float),
                             out_dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, m->stream));
    }
    HIPCK(hipStreamSynchronize(m->stream));
    m->ill_valid = m->ill_stride > 0;
    return OTFWI_OK;
}

template <typename R>
int gradient_impl(otfwi_model *m, int n, const int *shot_src, int kind, double *mis_out,
                  float *grad_out, bool out_dev, float *rec_out)
{
    int B = 0;
    RCK(grad_prepare(m, n, kind, &B));
    for (int s0 = 0; s0 < n; s0 += B)
        RCK(grad_batch_device<R>(m, s0, std::min(B, n - s0), shot_src, kind, mis_out, rec_out));
    return grad_complete(m, n, grad_out, out_dev);
}

// Two-step arguments of a fused loop phase for one stream group: for launch
// k, tab[j][k * nm ..] the problems it covers (mode j's arguments, in model
// order) and wg[k * (nm + 1) ..] the prefix sums of their work-group counts;
// dev holds both tables and then the prefix table on the device.
struct FusedGroup {
    std::vector<Pair2Args<float>> tab[2];   // [0] forward or reconstruction, [1] lambda
    std::vector<int> wg, cnt;
    std::vector<long long> nwg;
    char *dev = nullptr;
};
struct FusedPhase {
    int nm = 0, K = 0, G = 1;
    FusedGroup grp[2];
};

// The fused two-step launches of a phase (modes[0] then, if given, modes[1]
// per pair index): group g's launches captured into one graph on stream g
// (the first model's stream, then its first extra stream), the graphs
// replayed between a fork and a join, so each group's launch tails are
// filled by the other group's work-groups.
int run_fused_phase(otfwi_model *m0, const FusedPhase &ph, int nmodes, const int *modes)
{
    hipStream_t st[2] = {m0->stream, m0->xs[0]};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    auto body = [&]() -> int {
        const size_t tb = (size_t)ph.K * ph.nm * sizeof(Pair2Args<float>);
        for (int g = 0; g < ph.G; g++) {
            const FusedGroup &fg = ph.grp[g];
            // a group no model has shots in (one-shot models forced into two groups)
            if (std::none_of(fg.cnt.begin(), fg.cnt.end(), [](int c) { return c > 0; })) continue;
            const Pair2Args<float> *dt[2] = {(const Pair2Args<float> *)fg.dev,
                                             (const Pair2Args<float> *)(fg.dev + tb)};
            const int *dwg = (const int *)(fg.dev + 2 * tb);
            hipGraph_t graph = nullptr;
            HIPCK(hipStreamBeginCapture(st[g], hipStreamCaptureModeThreadLocal));
            hipError_t e = hipSuccess;
            for (int k = 0; k < ph.K && e == hipSuccess; k++)
                for (int j = 0; j < nmodes && e == hipSuccess; j++)
                    if (fg.cnt[k] > 0)
                        e = launch_pair2_multi(dt[j] + (size_t)k * ph.nm,
                                               dwg + (size_t)k * (ph.nm + 1), fg.cnt[k],
                                               fg.nwg[k], modes[j], st[g]);
            const hipError_t ec = hipStreamEndCapture(st[g], &graph);
            if (e != hipSuccess || ec != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                HIPCK(e);
                HIPCK(ec);
            }
            e = hipGraphInstantiate(&exec[g], graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            HIPCK(e);
        }
        if (ph.G > 1) RCK(fork_streams(m0, ph.G));
        for (int g = 0; g < ph.G; g++)
            if (exec[g]) HIPCK(hipGraphLaunch(exec[g], st[g]));
        if (ph.G > 1) RCK(join_streams(m0, ph.G));
        return OTFWI_OK;
    };
    const int rc = body();
    for (int g = 0; g < 2; g++)
        if (st[g]) (void)hipStreamSynchronize(st[g]);
    for (int g = 0; g < 2; g++)
        if (exec[g]) (void)hipGraphExecDestroy(exec[g]);
    return rc;
}

// Gradients of several FP32 models with every two-step launch covering all of
// them: per pair index one pair2_multi launch over the models whose loops
// still have that pair (uneven grids, shot counts and step counts) per stream
// group of shots, each group's launches of a phase captured into one graph.
// Each model's single steps, misfit and reductions run on its own stream,
// joined to the fused phases with events.  Results are bitwise those of
// separate gradient calls (the same kernels, cells and steps).
int fused_gradient_impl(int nm, otfwi_model *const *ms, const int *ns, const int *const *shot_src,
                        const int *kinds, double *const *mis_out, float *const *grad_out)
{
    using R = float;
    for (int i = 0; i < nm; i++) {
        otfwi_model *m = ms[i];
        const int n = ns[i];
        if (m->prec != 4) return fail(OTFWI_E_ARG, "fused gradient: model %d is not FP32", i);
        int B = 0;
        RCK(grad_prepare(m, n, kinds[i], &B));
        if (B != n)
            return fail(OTFWI_E_ARG, "fused gradient: model %d's %d shots exceed its batch limit %d",
                        i, n, B);
        if (m->ill_stride > 0)
            return fail(OTFWI_E_ARG, "fused gradient: model %d collects illumination (one-step "
                        "forward loops); switch it off", i);
        // every model runs two steps per launch whatever its own launch size
        // (the fused launches are the sum of the models'), unless one step per
        // launch was asked for
        const size_t plane = (size_t)m->zp * m->ld;
        const bool pairs_ok = m->evrow && m->ringtab && m->fwd_steps != 1 && m->adj_steps != 1 &&
                              m->tb_steps[0] <= 1 && m->tb_steps[1] <= 1 &&
                              m->tb_shots[0] <= 0 && m->tb_shots[1] <= 0 &&
                              pair2_supported((long long)plane, (long long)m->b.cap * (long long)plane,
                                              (int)sizeof(R));
        if (!pairs_ok || m->pair_nw_built != F2_MULTI_NW)
            return fail(OTFWI_E_ARG, "fused gradient: model %d cannot run two steps per launch "
                        "of %d threads (grid, batch or settings); otfwi_gradient_multi runs it",
                        i, F2_MULTI_NW);
        RCK(ensure_lap<R>(m));
        RCK(upload_shot_src(m, n, shot_src ? shot_src[i] : nullptr));
        RCK(upload_obs_batch<R>(m, 0, n));
    }
    // argument tables of both phases, split into G stream groups of each
    // model's shots (two when half of the first launch still makes 2
    // work-groups per CU, as pair_split decides for one model)
    FusedPhase tf, ta;
    tf.nm = ta.nm = nm;
    std::vector<int> nlf(nm), npf(nm), nla(nm), npa(nm);
    long long half_wg = 0;
    for (int i = 0; i < nm; i++) {
        fwd_schedule(1, ms[i]->T, nlf[i], npf[i]);
        adj_schedule(1, ms[i]->T, nla[i], npa[i]);
        tf.K = std::max(tf.K, npf[i]);
        ta.K = std::max(ta.K, npa[i]);
        half_wg += pair2_groups(ms[i]->X, ms[i]->Z, F2_MULTI_NW, ns[i] / 2);
    }
    // the first model's pair-stream setting decides when it is not automatic
    const int ps = ms[0]->pair_streams;
    const int G = !ms[0]->xs[0] || ps == 1 ? 1 :
                  ps >= 2 ? 2 : (half_wg >= 2LL * std::max(ms[0]->n_cu, 1) ? 2 : 1);
    tf.G = ta.G = G;
    for (FusedPhase *ph : {&tf, &ta}) {
        const bool fwd = ph == &tf;
        for (int g = 0; g < G; g++) {
            FusedGroup &fg = ph->grp[g];
            fg.tab[0].assign((size_t)ph->K * nm, Pair2Args<float>{});
            fg.tab[1].assign((size_t)ph->K * nm, Pair2Args<float>{});
            fg.wg.assign((size_t)ph->K * (nm + 1), 0);
            fg.cnt.assign(ph->K, 0);
            fg.nwg.assign(ph->K, 0);
            for (int k = 0; k < ph->K; k++) {
                int c = 0;
                long long w = 0;
                for (int i = 0; i < nm; i++) {
                    otfwi_model *m = ms[i];
                    const int n = ns[i];
                    int g0, g1;
                    group_range(n, g, G, g0, g1);
                    if (k >= (fwd ? npf[i] : npa[i]) || g1 <= g0) continue;
                    if (fwd) {
                        Pair2Args<R> q = pair_fwd_args<R>(m, n, fwd_args<R>(m, n));
                        fwd_pair_slots<R>(m, q, k, 1 + nlf[i] + 2 * k);
                        q = pair_group<R>(m, q, g0, g1 - g0);
                        if (!pair2_args_ok(q, 0))
                            return fail(OTFWI_E_STATE, "fused gradient: model %d forward arguments", i);
                        fg.tab[0][(size_t)k * nm + c] = q;
                    } else {
                        Pair2Args<R> u = pair_rec_args<R>(m, n), l = pair_lam_args<R>(m, n);
                        adj_pair_slots<R>(m, u, l, k, 1 + nla[i] + 2 * k);
                        u = pair_group<R>(m, u, g0, g1 - g0);
                        l = pair_group<R>(m, l, g0, g1 - g0);
                        if (!pair2_args_ok(u, 2) || !pair2_args_ok(l, 1))
                            return fail(OTFWI_E_STATE, "fused gradient: model %d adjoint arguments", i);
                        fg.tab[0][(size_t)k * nm + c] = u;
                        fg.tab[1][(size_t)k * nm + c] = l;
                    }
                    fg.wg[(size_t)k * (nm + 1) + c] = (int)w;
                    w += pair2_groups(m->X, m->Z, F2_MULTI_NW, g1 - g0);
                    if (w > 0x7FFFFFFFLL)
                        return fail(OTFWI_E_ARG, "fused gradient: too many work-groups per launch");
                    c++;
                }
                fg.wg[(size_t)k * (nm + 1) + c] = (int)w;
                fg.cnt[k] = c;
                fg.nwg[k] = w;
            }
        }
    }
    // one device allocation per phase and group: both argument tables, then the prefix table
    int rc = OTFWI_OK;
    hipStream_t main = ms[0]->stream;
    auto upload = [&](FusedPhase &ph) -> int {
        const size_t tb = (size_t)ph.K * nm * sizeof(Pair2Args<float>);
        for (int g = 0; g < ph.G; g++) {
            FusedGroup &fg = ph.grp[g];
            const size_t bytes = 2 * tb + fg.wg.size() * sizeof(int);
            if (bytes == 0) continue;
            RCK(dalloc(&fg.dev, bytes));
            HIPCK(hipMemcpyAsync(fg.dev, fg.tab[0].data(), tb, hipMemcpyHostToDevice, main));
            HIPCK(hipMemcpyAsync(fg.dev + tb, fg.tab[1].data(), tb, hipMemcpyHostToDevice, main));
            HIPCK(hipMemcpyAsync(fg.dev + 2 * tb, fg.wg.data(), fg.wg.size() * sizeof(int),
                                 hipMemcpyHostToDevice, main));
        }
        HIPCK(hipStreamSynchronize(main));
        return OTFWI_OK;
    };
    auto body = [&]() -> int {
        RCK(upload(tf));
        RCK(upload(ta));
        // forward: memsets and leading single steps on each model's stream
        for (int i = 0; i < nm; i++) {
            otfwi_model *m = ms[i];
            HIPCK(hipEventRecord(m->ev[0], m->stream));
            RCK(run_forward<R>(m, ns[i], true, false, LOOP_LEAD));
            HIPCK(hipEventRecord(m->evfj[1], m->stream));
            HIPCK(hipStreamWaitEvent(main, m->evfj[1], 0));
        }
        const int mf[1] = {0}, ma[2] = {2, 1};
        if (tf.K > 0) RCK(run_fused_phase(ms[0], tf, 1, mf));
        HIPCK(hipEventRecord(ms[0]->evfj[2], main));
        // last forward step, misfit, adjoint memsets and leading steps per model
        for (int i = 0; i < nm; i++) {
            otfwi_model *m = ms[i];
            HIPCK(hipStreamWaitEvent(m->stream, ms[0]->evfj[2], 0));
            RCK(run_forward<R>(m, ns[i], true, false, LOOP_TRAIL));
            HIPCK(hipEventRecord(m->ev[1], m->stream));
            RCK(batch_misfit<R>(m, ns[i], kinds[i]));
            HIPCK(hipEventRecord(m->ev[2], m->stream));
            RCK(run_adjoint<R>(m, ns[i], LOOP_LEAD));
            HIPCK(hipEventRecord(m->evfj[1], m->stream));
            HIPCK(hipStreamWaitEvent(main, m->evfj[1], 0));
        }
        if (ta.K > 0) RCK(run_fused_phase(ms[0], ta, 2, ma));
        HIPCK(hipEventRecord(ms[0]->evfj[2], main));
        for (int i = 0; i < nm; i++) {
            otfwi_model *m = ms[i];
            HIPCK(hipStreamWaitEvent(m->stream, ms[0]->evfj[2], 0));
            HIPCK(hipEventRecord(m->ev[3], m->stream));
            RCK(grad_batch_finish<R>(m, 0, ns[i], kinds[i], mis_out ? mis_out[i] : nullptr));
            RCK(grad_complete(m, ns[i], grad_out ? grad_out[i] : nullptr, false));
        }
        return OTFWI_OK;
    };
    std::vector<int> steps_set(2 * nm);
    for (int i = 0; i < nm; i++) {
        steps_set[2 * i] = ms[i]->fwd_steps;
        steps_set[2 * i + 1] = ms[i]->adj_steps;
        ms[i]->fwd_steps = ms[i]->adj_steps = 2;
    }
    rc = body();
    for (int i = 0; i < nm; i++) {
        ms[i]->fwd_steps = steps_set[2 * i];
        ms[i]->adj_steps = steps_set[2 * i + 1];
    }
    // every stream idle before the tables are released
    for (int i = 0; i < nm; i++) (void)hipStreamSynchronize(ms[i]->stream);
    for (int k = 0; k < 3; k++)
        if (ms[0]->xs[k]) (void)hipStreamSynchronize(ms[0]->xs[k]);
    for (FusedPhase *ph : {&tf, &ta})
        for (int g = 0; g < 2; g++) dfree(ph->grp[g].dev);
    return rc;
}

// Misfit of the current model without the adjoint: forward runs of the shots
// (no ring save), device misfit, per-shot FP64 sums into mis_out.
template <typename R>
int objective_impl(otfwi_model *m, int n, const int *shot_src, int kind, double *mis_out)
{
    if (!misfit_kind_ok(kind)) return fail(OTFWI_E_ARG, "unknown misfit kind %d", kind);
    if (kind == OTFWI_MISFIT_ENV) RCK(ensure_hilb(m));
    if (n <= 0) return fail(OTFWI_E_ARG, "n_shots must be positive");
    if (m->n_rx <= 0) return fail(OTFWI_E_STATE, "receivers not set");
    if (w2_family(kind) && m->T < 2) return fail(OTFWI_E_ARG, "W2 needs T >= 2");
    RCK(ready(m, false));
    const int B = batch_size(m, n);
    RCK(ensure_batch(m, B, true, false));
    Batch &b = m->b;
    for (int s0 = 0; s0 < n; s0 += B) {
        const int nb = std::min(B, n - s0);
        RCK(upload_shot_src(m, nb, shot_src ? shot_src + s0 : nullptr));
        RCK(upload_obs_batch<R>(m, s0, nb));
        RCK(run_forward<R>(m, nb, false));
        const int ntr = nb * m->n_rx;
        int *flags = misfit_flags(kind, b.scratch, ntr, m->T);
        if (w2_family(kind)) HIPCK(hipMemsetAsync(flags, 0, ntr * sizeof(int), m->stream));
        HIPCK((launch_misfit<R, float>(kind, (const R *)b.rec, b.obs, (R *)b.adj, nullptr,
                                       b.mis_tr, ntr, m->T, b.scratch, m->stream, m->hilb)));
        HIPCK(launch_reduce_misfit(b.mis_tr, nb, m->n_rx, b.mis_shot, m->stream));
        std::vector<double> hmis(nb);
        HIPCK(hipMemcpyAsync(hmis.data(), b.mis_shot, nb * sizeof(double), hipMemcpyDeviceToHost,
                             m->stream));
        std::vector<int> fl(w2_family(kind) ? ntr : 0);
        if (w2_family(kind))
            HIPCK(hipMemcpyAsync(fl.data(), flags, ntr * sizeof(int), hipMemcpyDeviceToHost,
                                 m->stream));
        HIPCK(hipStreamSynchronize(m->stream));
        RCK(check_w2_flags(m, fl, s0));
        for (int i = 0; i < nb; i++) mis_out[s0 + i] = hmis[i];
    }
    return OTFWI_OK;
}

template <typename R>
int forward_for_grad_impl(otfwi_model *m, int n, const int *shot_src, float *rec_out)
{
    if (n <= 0) return fail(OTFWI_E_ARG, "n_shots must be positive");
    RCK(ready(m, true));
    if (m->max_batch > 0 && n > m->max_batch)
        return fail(OTFWI_E_ARG, "two-stage gradient takes at most max_batch=%d shots per call",
                    m->max_batch);
    RCK(ensure_batch(m, n, true));
    RCK(upload_shot_src(m, n, shot_src));
    RCK(run_forward<R>(m, n, true));
    if (rec_out) RCK(copy_rec_out<R>(m, n, rec_out));
    HIPCK(hipStreamSynchronize(m->stream));
    m->pend_n = n;
    return OTFWI_OK;
}

template <typename R>
int adjoint_from_host_impl(otfwi_model *m, const float *adj, float *grad_out)
{
    const int n = m->pend_n;
    if (n <= 0) return fail(OTFWI_E_STATE, "no pending forward run (call otfwi_forward_for_grad)");
    if (!adj) return fail(OTFWI_E_ARG, "adjoint source is NULL");
    Batch &b = m->b;
    const size_t cnt = (size_t)n * m->n_rx * m->T;
    std::vector<R> tmp(cnt);
    for (size_t i = 0; i < cnt; i++) tmp[i] = (R)adj[i];
    HIPCK(hipMemcpyAsync(b.adj, tmp.data(), cnt * sizeof(R), hipMemcpyHostToDevice, m->stream));
    HIPCK(launch_jcum<R>((const R *)b.adj, (R *)b.jcum, n * m->n_rx, m->T, m->stream));
    RCK(run_adjoint<R>(m, n));
    RCK(ensure_gacc(m));
    const size_t plane = (size_t)m->zp * m->ld;
    HIPCK(launch_accum_grad<R>((const R *)b.grad, n, m->X, m->Z, m->ld, (long long)plane, m->gacc,
                               m->stream));
    const int cells = m->X * m->Z;
    HIPCK(launch_finalize_grad(m->gacc, cells, 1.0 / n, m->gout, m->stream));
    if (grad_out)
        HIPCK(hipMemcpyAsync(grad_out, m->gout, cells * sizeof(float), hipMemcpyDeviceToHost,
                             m->stream));
    HIPCK(hipStreamSynchronize(m->stream));
    m->pend_n = 0;
    return OTFWI_OK;
}

// Padded receiver map for the exact adjoint: receiver index at its cell,
// -1 elsewhere (receivers outside the updated box record 0 and get none).
int build_rxmap(otfwi_model *m)
{
    const size_t plane = (size_t)m->zp * m->ld;
    std::vector<int> map(plane, -1);
    for (int r = 0; r < m->n_rx; r++) {
        const int x = m->rxx[r], z = m->rxz[r];
        if (x < BOX || x >= m->X - BOX || z < BOX || z >= m->Z - BOX) continue;
        int &slot = map[(size_t)(z + ZOFF) * m->ld + x + XOFF];
        if (slot >= 0)
            return fail(OTFWI_E_ARG, "receivers %d and %d share cell (%d, %d)", slot, r, x, z);
        slot = r;
    }
    dfree(m->rxmap);
    RCK(dalloc(&m->rxmap, plane));
    HIPCK(hipMemcpy(m->rxmap, map.data(), plane * sizeof(int), hipMemcpyHostToDevice));
    return OTFWI_OK;
}

int ensure_xbuf(otfwi_model *m, size_t snap_bytes, size_t seg_bytes)
{
    if (m->xsnap_bytes < snap_bytes) {
        dfree(m->xsnap);
        m->xsnap_bytes = 0;
        RCK(dalloc((char **)&m->xsnap, snap_bytes));
        m->xsnap_bytes = snap_bytes;
    }
    if (m->xseg_bytes < seg_bytes) {
        dfree(m->xseg);
        m->xseg_bytes = 0;
        RCK(dalloc((char **)&m->xseg, seg_bytes));
        m->xseg_bytes = seg_bytes;
    }
    // padding and the never-updated ring of the segment A fields stay 0
    HIPCK(hipMemsetAsync(m->xseg, 0, m->xseg_bytes, m->stream));
    return OTFWI_OK;
}

// Gradient of J = mean over shots of the misfit value (L2: sum (sim-obs)^2;
// W2: the reference W2 value) by the exact discrete adjoint, with the forward
// A fields recomputed segment by segment from snapshots taken every k steps
// (k <= 0: round(sqrt(T-1))).  Memory per shot: ceil((T-1)/k) snapshots of 4
// planes + k planes of (Ax, Az), instead of T-1 of them.  Returns dJ/d(c^2)
// (mean over shots) in FP32 and/or FP64.
template <typename R>
int exact_gradient_impl(otfwi_model *m, int n, const int *shot_src, int kind, int k_snap,
                        double *mis_out, float *grad_out, double *grad64_out, bool out_dev = false)
{
    if (kind != OTFWI_MISFIT_L2 && kind != OTFWI_MISFIT_W2)
        return fail(OTFWI_E_ARG, "misfit kind %d: the exact gradient supports L2 and W2", kind);
    if (n <= 0) return fail(OTFWI_E_ARG, "n_shots must be positive");
    if (m->n_rx <= 0) return fail(OTFWI_E_STATE, "receivers not set");
    if (m->T < 2) return fail(OTFWI_E_ARG, "exact gradient needs T >= 2");
    RCK(ready(m, false));
    const int T = m->T;
    const int k = k_snap > 0 ? std::min(k_snap, T - 1)
                             : std::max(1, (int)std::lround(std::sqrt((double)(T - 1))));
    const int nseg = ceil_div(T - 1, k);
    const int B = batch_size(m, n);
    RCK(ensure_batch(m, B, true, false));
    RCK(ensure_gacc(m));
    RCK(build_rxmap(m));
    Batch &b = m->b;
    const size_t es = sizeof(R), plane = (size_t)m->zp * m->ld;
    const size_t bp = plane * B;     // one field of the batch
    RCK(ensure_xbuf(m, (size_t)nseg * 4 * bp * es, (size_t)k * 2 * bp * es));
    hipStream_t st = m->stream;
    R *snap = (R *)m->xsnap, *seg = (R *)m->xseg;
    auto snap_f = [&](int j, int f) { return snap + ((size_t)j * 4 + f) * bp; };
    auto seg_ax = [&](int i) { return seg + (size_t)i * 2 * bp; };
    auto seg_az = [&](int i) { return seg + ((size_t)i * 2 + 1) * bp; };
    const int pw = persistent_wgs(m, false);
    m->t_fwd_ms = m->t_mis_ms = m->t_adj_ms = 0;
    m->n_fwd_launch = m->n_adj_launch = 0;
    for (int s0 = 0; s0 < n; s0 += B) {
        const int nb = std::min(B, n - s0);
        const size_t fb = plane * nb * es;   // bytes of one field of this batch
        RCK(upload_shot_src(m, nb, shot_src ? shot_src + s0 : nullptr));
        RCK(upload_obs_batch<R>(m, s0, nb));
        for (int i = 0; i < 2; i++)
            for (void *p : {b.P[i], b.Ax[i], b.Az[i], b.Px[i]}) HIPCK(hipMemsetAsync(p, 0, fb, st));
        HIPCK(hipMemsetAsync(b.rec, 0, (size_t)nb * m->n_rx * T * es, st));
        FwdArgs<R> fa = fwd_args<R>(m, nb);
        fa.ring = nullptr;
        HIPCK(hipEventRecord(m->ev[0], st));
        // phase 1: forward with a snapshot of the state at t-1 before every
        // segment start t = 1 + j k
        for (int t = 1; t < T; t++) {
            const int p = (t - 1) & 1, c = t & 1;
            if ((t - 1) % k == 0) {
                const int j = (t - 1) / k;
                HIPCK(hipMemcpyAsync(snap_f(j, 0), b.P[p], fb, hipMemcpyDeviceToDevice, st));
                HIPCK(hipMemcpyAsync(snap_f(j, 1), b.Px[p], fb, hipMemcpyDeviceToDevice, st));
                HIPCK(hipMemcpyAsync(snap_f(j, 2), b.Ax[p], fb, hipMemcpyDeviceToDevice, st));
                HIPCK(hipMemcpyAsync(snap_f(j, 3), b.Az[p], fb, hipMemcpyDeviceToDevice, st));
            }
            fa.t = t;
            fa.P0 = (const R *)b.P[p]; fa.P1 = (R *)b.P[c];
            fa.Ax0 = (const R *)b.Ax[p]; fa.Ax1 = (R *)b.Ax[c];
            fa.Az0 = (const R *)b.Az[p]; fa.Az1 = (R *)b.Az[c];
            fa.Px0 = (const R *)b.Px[p]; fa.Px1 = (R *)b.Px[c];
            HIPCK(launch_fwd_step<R>(fa, m->th_fwd, pw, st));
        }
        HIPCK(hipEventRecord(m->ev[1], st));
        // misfit value and its exact derivative rho = dJ/d(trace)
        const int ntr = nb * m->n_rx;
        int *flags = reinterpret_cast<int *>(b.scratch + (size_t)ntr * 3 * T);
        HIPCK(hipMemsetAsync(flags, 0, ntr * sizeof(int), st));
        HIPCK((launch_misfit_exact<R, float>(kind, (const R *)b.rec, b.obs, (R *)b.adj, b.mis_tr,
                                             ntr, T, b.scratch, flags, st)));
        HIPCK(launch_reduce_misfit(b.mis_tr, nb, m->n_rx, b.mis_shot, st));
        HIPCK(hipEventRecord(m->ev[2], st));
        // phase 2: segments from the last; recompute A, then run the adjoint back
        for (int i = 0; i < 2; i++)
            for (void *p : {b.LP[i], b.LPx[i], b.LAx[i], b.LAz[i]})
                HIPCK(hipMemsetAsync(p, 0, fb, st));
        HIPCK(hipMemsetAsync(b.grad, 0, fb, st));
        XAdjArgs<R> xa{};
        xa.X = m->X; xa.Z = m->Z; xa.ld = m->ld; xa.nshots = nb; xa.n_rx = m->n_rx; xa.T = T;
        xa.sstride = (long long)plane;
        xa.grad = (R *)b.grad;
        xa.cm = (const R *)m->cm; xa.kx = (const R *)m->kx; xa.kz = (const R *)m->kz;
        xa.cA = (R)m->dt / (R)m->dx;
        xa.rxmap = m->rxmap;
        xa.rho = (const R *)b.adj;
        int cur = 0;   // LP/LPx[cur] hold u, w of step t+1; LAx/LAz[cur] hold a_t, b_t
        for (int j = nseg - 1; j >= 0; j--) {
            const int ts = 1 + j * k, te = std::min(ts + k, T);
            const int p0 = (ts - 1) & 1;
            HIPCK(hipMemcpyAsync(b.P[p0], snap_f(j, 0), fb, hipMemcpyDeviceToDevice, st));
            HIPCK(hipMemcpyAsync(b.Px[p0], snap_f(j, 1), fb, hipMemcpyDeviceToDevice, st));
            HIPCK(hipMemcpyAsync(seg_ax(0), snap_f(j, 2), fb, hipMemcpyDeviceToDevice, st));
            HIPCK(hipMemcpyAsync(seg_az(0), snap_f(j, 3), fb, hipMemcpyDeviceToDevice, st));
            for (int t = ts; t <= te - 2; t++) {
                const int p = (t - 1) & 1, c = t & 1;
                fa.t = t;
                fa.P0 = (const R *)b.P[p]; fa.P1 = (R *)b.P[c];
                fa.Px0 = (const R *)b.Px[p]; fa.Px1 = (R *)b.Px[c];
                fa.Ax0 = seg_ax(t - ts); fa.Ax1 = seg_ax(t - ts + 1);
                fa.Az0 = seg_az(t - ts); fa.Az1 = seg_az(t - ts + 1);
                HIPCK(launch_fwd_step<R>(fa, m->th_fwd, pw, st));
                m->n_fwd_launch++;
            }
            for (int t = te - 1; t >= ts; t--) {
                xa.t = t;
                xa.u0 = (const R *)b.LP[cur]; xa.w0 = (const R *)b.LPx[cur];
                xa.a = (const R *)b.LAx[cur]; xa.b = (const R *)b.LAz[cur];
                xa.u1 = (R *)b.LP[1 - cur]; xa.w1 = (R *)b.LPx[1 - cur];
                xa.a1 = (R *)b.LAx[1 - cur]; xa.b1 = (R *)b.LAz[1 - cur];
                xa.Ax = seg_ax(t - ts); xa.Az = seg_az(t - ts);
                HIPCK(launch_xadj_step<R>(xa, st));
                m->n_adj_launch++;
                cur = 1 - cur;
            }
        }
        HIPCK(hipEventRecord(m->ev[3], st));
        HIPCK(launch_accum_grad<R>((const R *)b.grad, nb, m->X, m->Z, m->ld, (long long)plane,
                                   m->gacc, st));
        std::vector<double> hmis(nb);
        std::vector<int> fl(ntr);
        HIPCK(hipMemcpyAsync(hmis.data(), b.mis_shot, nb * sizeof(double), hipMemcpyDeviceToHost,
                             st));
        HIPCK(hipMemcpyAsync(fl.data(), flags, ntr * sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCK(hipStreamSynchronize(st));
        float ms[3];
        HIPCK(hipEventElapsedTime(&ms[0], m->ev[0], m->ev[1]));
        HIPCK(hipEventElapsedTime(&ms[1], m->ev[1], m->ev[2]));
        HIPCK(hipEventElapsedTime(&ms[2], m->ev[2], m->ev[3]));
        m->t_fwd_ms += ms[0];
        m->t_mis_ms += ms[1];
        m->t_adj_ms += ms[2];
        m->n_fwd_launch += T - 1;
        for (size_t i = 0; i < fl.size(); i++)
            if (fl[i] == 3)
                return fail(OTFWI_E_NUMERIC,
                            "exact W2: the CDF of simulated trace %zu of shot %d is not monotone "
                            "(sim + c < 0 somewhere)", i % m->n_rx, s0 + (int)(i / m->n_rx));
        RCK(check_w2_flags(m, fl, s0));
        if (mis_out)
            for (int i = 0; i < nb; i++) mis_out[s0 + i] = hmis[i];
    }
    const int cells = m->X * m->Z;
    const double scale = ((double)m->dt / (double)m->dx) / n;   // m = c^2 dt/dx
    if (grad_out) {
        HIPCK(launch_finalize_grad(m->gacc, cells, scale, m->gout, st));
        HIPCK(hipMemcpyAsync(grad_out, m->gout, cells * sizeof(float),
                             out_dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
    }
    if (grad64_out) {
        std::vector<double> acc(cells);
        HIPCK(hipMemcpyAsync(acc.data(), m->gacc, cells * sizeof(double), hipMemcpyDeviceToHost,
                             st));
        HIPCK(hipStreamSynchronize(st));
        for (int i = 0; i < cells; i++) grad64_out[i] = acc[i] * scale;
    }
    HIPCK(hipStreamSynchronize(st));
    return OTFWI_OK;
}

// Per-launch device time (HIP events on the model's stream, mean of reps
// launches after one warm-up launch) of the step kernels on the current
// batch of n shots: ms[0] one forward step, ms[1] one adjoint step, ms[2] a
// forward pair, ms[3] a reconstruction pair, ms[4] a lambda pair.  The
// fields hold whatever the batch buffers hold (zeros after a memset); the
// kernels' cost does not depend on the values.
template <typename R>
int kernel_times_impl(otfwi_model *m, int n, int reps, float *ms)
{
    if (n <= 0 || reps <= 0) return fail(OTFWI_E_ARG, "kernel times: n_shots and reps must be > 0");
    if (m->T < 4) return fail(OTFWI_E_ARG, "kernel times: need T >= 4");
    RCK(ready(m, true));
    RCK(ensure_batch(m, n, true));
    RCK(ensure_lap<R>(m));
    RCK(upload_shot_src(m, n, nullptr));
    Batch &b = m->b;
    hipStream_t st = m->stream;
    const bool have_pairs = m->evrow && m->ringtab &&
                            pair2_supported((long long)m->zp * m->ld, (long long)m->b.cap * m->zp * m->ld, (int)sizeof(R));
    FwdArgs<R> fa = fwd_args<R>(m, n);
    fa.t = 1;
    fa.P0 = (const R *)b.P[0]; fa.P1 = (R *)b.P[1];
    fa.Ax0 = (const R *)b.Ax[0]; fa.Ax1 = (R *)b.Ax[1];
    fa.Az0 = (const R *)b.Az[0]; fa.Az1 = (R *)b.Az[1];
    fa.Px0 = (const R *)b.Px[0]; fa.Px1 = (R *)b.Px[1];
    AdjArgs<R> aa = adj_args<R>(m, n);
    aa.t = 1;
    aa.LP0 = (const R *)b.LP[0]; aa.LP1 = (R *)b.LP[1];
    aa.LAx0 = (const R *)b.LAx[0]; aa.LAx1 = (R *)b.LAx[1];
    aa.LAz0 = (const R *)b.LAz[0]; aa.LAz1 = (R *)b.LAz[1];
    aa.LPx0 = (const R *)b.LPx[0]; aa.LPx1 = (R *)b.LPx[1];
    aa.U0 = (const R *)b.P[0]; aa.U1 = (R *)b.P[1];
    aa.UAx0 = (const R *)b.Ax[0]; aa.UAx1 = (R *)b.Ax[1];
    aa.UAz0 = (const R *)b.Az[0]; aa.UAz1 = (R *)b.Az[1];
    Pair2Args<R> pf = pair_fwd_args<R>(m, n, fa), pu = pair_rec_args<R>(m, n),
                 pl = pair_lam_args<R>(m, n);
    for (Pair2Args<R> *q : {&pf, &pu}) {
        q->t = 1;
        q->P0 = (const R *)b.P[0]; q->P1 = (R *)b.P[1];
        q->Ax0 = (const R *)b.Ax[0]; q->Ax1 = (R *)b.Ax[1];
        q->Az0 = (const R *)b.Az[0]; q->Az1 = (R *)b.Az[1];
        q->Px0 = (const R *)b.Px[0]; q->Px1 = (R *)b.Px[1];
    }
    pl.t = 1;
    pl.P0 = (const R *)b.LP[0]; pl.P1 = (R *)b.LP[1];
    pl.Ax0 = (const R *)b.LAx[0]; pl.Ax1 = (R *)b.LAx[1];
    pl.Az0 = (const R *)b.LAz[0]; pl.Az1 = (R *)b.LAz[1];
    pl.Px0 = (const R *)b.LPx[0]; pl.Px1 = (R *)b.LPx[1];
    const int pwf = persistent_wgs(m, false), pwa = persistent_wgs(m, true);
    for (int k = 0; k < 5; k++) {
        ms[k] = -1.0f;
        if (k >= 2 && !have_pairs) continue;
        auto launch = [&]() -> hipError_t {
            switch (k) {
            case 0: return launch_fwd_step<R>(fa, m->th_fwd, pwf, st);
            case 1: return launch_adj_step<R>(aa, m->th_adj, pwa, st);
            case 2: return launch_pair2<R>(pf, 0, st);
            case 3: return launch_pair2<R>(pu, 2, st);
            default: return launch_pair2<R>(pl, 1, st);
            }
        };
        HIPCK(launch());
        HIPCK(hipEventRecord(m->ev[0], st));
        for (int r = 0; r < reps; r++) HIPCK(launch());
        HIPCK(hipEventRecord(m->ev[1], st));
        HIPCK(hipEventSynchronize(m->ev[1]));
        float t = 0;
        HIPCK(hipEventElapsedTime(&t, m->ev[0], m->ev[1]));
        ms[k] = t / reps;
    }
    return OTFWI_OK;
}

// Times `steps` launches of each tile height / launch style of the forward
// and adjoint step kernels on the live batch of n shots and keeps the
// fastest, then picks the steps per launch (1 or 2) of both loops by time
// per step.  out[0..3] = th_fwd, persistent_fwd, th_adj, persistent_adj;
// ms[0], ms[1] = best average launch time (ms) of the one-step kernels.
template <typename R>
int autotune_impl(otfwi_model *m, int n, int steps, int *out, float *ms)
{
    if (n <= 0 || steps <= 0) return fail(OTFWI_E_ARG, "autotune: n_shots and steps must be > 0");
    steps = std::min(steps, m->T - 1);
    RCK(ready(m, true));
    RCK(ensure_batch(m, n, true));
    RCK(upload_shot_src(m, n, nullptr));
    Batch &b = m->b;
    const size_t es = sizeof(R), plane = (size_t)m->zp * m->ld;
    for (int i = 0; i < 2; i++)
        for (void *p : {b.P[i], b.Ax[i], b.Az[i], b.Px[i], b.LP[i], b.LAx[i], b.LAz[i], b.LPx[i]})
            HIPCK(hipMemsetAsync(p, 0, plane * n * es, m->stream));
    HIPCK(hipMemsetAsync(b.grad, 0, plane * n * es, m->stream));
    HIPCK(hipMemsetAsync(b.jcum, 0, (size_t)n * m->n_rx * m->T * es, m->stream));
    HIPCK(hipMemsetAsync(b.ring, 0, (size_t)n * m->T * m->n_ring * es, m->stream));
    const bool graph = m->use_graph;
    m->use_graph = false;
    const int steps_saved[2] = {m->fwd_steps, m->adj_steps};
    m->fwd_steps = m->adj_steps = 1;
    const int ths[3] = {16, 32, 64};
    int best[4] = {m->th_fwd, m->pers_fwd, m->th_adj, m->pers_adj};
    float bms[2] = {1e30f, 1e30f};
    int rc = OTFWI_OK;
    for (int adj = 0; adj < 2 && rc == OTFWI_OK; adj++)
        for (int th : ths)
            for (int pers = 0; pers < (m->prec == 8 ? 1 : 2) && rc == OTFWI_OK; pers++) {
                if (m->prec == 8 && (th == 64 || (adj && th != 16))) continue;
                if (adj) { m->th_adj = th; m->pers_adj = pers; }
                else { m->th_fwd = th; m->pers_fwd = pers; }
                if ((rc = ready(m, true)) != OTFWI_OK) break;
                auto launch = [&](int k) {
                    return adj ? adj_loop<R>(m, n, 1, 1 + k) : fwd_loop<R>(m, n, true, 1, 1 + k);
                };
                if ((rc = launch(std::min(2, steps))) != OTFWI_OK) break;
                float t = 0;
                if (hipEventRecord(m->ev[0], m->stream) != hipSuccess ||
                    (rc = launch(steps)) != OTFWI_OK ||
                    hipEventRecord(m->ev[1], m->stream) != hipSuccess ||
                    hipEventSynchronize(m->ev[1]) != hipSuccess ||
                    hipEventElapsedTime(&t, m->ev[0], m->ev[1]) != hipSuccess) {
                    if (rc == OTFWI_OK) rc = fail(OTFWI_E_HIP, "autotune timing failed");
                    break;
                }
                t /= steps;
                if (t < bms[adj]) {
                    bms[adj] = t;
                    best[2 * adj] = th;
                    best[2 * adj + 1] = pers;
                }
            }
    m->th_fwd = best[0]; m->pers_fwd = best[1];
    m->th_adj = best[2]; m->pers_adj = best[3];
    m->fwd_steps = steps_saved[0];
    m->adj_steps = steps_saved[1];
    if (rc == OTFWI_OK) rc = ready(m, true);
    // steps per launch: two-step launches when their time per step is lower
    if (rc == OTFWI_OK) {
        float kt[5] = {};
        rc = kernel_times_impl<R>(m, n, 5, kt);
        if (rc == OTFWI_OK) {
            m->fwd_steps = kt[2] / 2 < kt[0] ? 2 : 1;
            m->adj_steps = (kt[3] + kt[4]) / 2 < kt[1] ? 2 : 1;
        }
    }
    m->use_graph = graph;
    RCK(rc);
    if (out) for (int i = 0; i < 4; i++) out[i] = best[i];
    if (ms) { ms[0] = bms[0]; ms[1] = bms[1]; }
    return OTFWI_OK;
}

#define DISPATCH(m, call) ((m)->prec == 8 ? call<double> : call<float>)

int check_model(const otfwi_model *m)
{
    if (!m) return fail(OTFWI_E_ARG, "model handle is NULL");
    return OTFWI_OK;
}

}  // namespace

// ===========================================================================
// Handle API
// ===========================================================================
extern "C" {

const char *otfwi_last_error(void) { return g_err.c_str(); }

int otfwi_device_info(char *buf, int len)
{
    int dev = 0;
    hipDeviceProp_t p;
    HIPCK(hipGetDevice(&dev));
    HIPCK(hipGetDeviceProperties(&p, dev));
    if (buf && len > 0)
        snprintf(buf, len, "%s | %s | %d CUs | %.1f GB", p.name, p.gcnArchName,
                 p.multiProcessorCount, p.totalGlobalMem / 1e9);
    return OTFWI_OK;
}

int otfwi_create(otfwi_model **out, int X, int Z, int T, int precision, const float *constvec)
{
    if (!out) return fail(OTFWI_E_ARG, "out is NULL");
    *out = nullptr;
    if (X < 2 * BOX + 1 || Z < 2 * BOX + 1 || T < 2)
        return fail(OTFWI_E_ARG, "grid %dx%d with T=%d is too small", X, Z, T);
    if ((long long)X * Z > (1LL << 30)) return fail(OTFWI_E_ARG, "grid too large");
    if (precision != 4 && precision != 8)
        return fail(OTFWI_E_ARG, "precision must be 4 (FP32) or 8 (FP64)");
    otfwi_model *m = new otfwi_model();
    m->X = X; m->Z = Z; m->T = T; m->prec = precision;
    m->ld = padded_ld(X);
    m->zp = padded_rows(Z);
    if (constvec) { m->dx = constvec[0]; m->dz = constvec[1]; m->dt = constvec[2]; }
    if (!(m->dx > 0) || !(m->dt > 0)) {
        delete m;
        return fail(OTFWI_E_ARG, "dx and dt must be positive");
    }
    if (precision == 8) { m->th_fwd = 16; m->th_adj = 16; }
    {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
            delete m;
            return fail(OTFWI_E_HIP, "no usable HIP device");
        }
        m->n_cu = prop.multiProcessorCount;
    }
    m->dxs_h.assign(X, 0.0f);
    m->dzs_h.assign(Z, 0.0f);
    m->kx_h.assign(X, 1.0f);
    m->kz_h.assign(Z, 1.0f);
    hipError_t e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
    for (int i = 0; i < 3 && e == hipSuccess; i++)
        e = hipStreamCreateWithFlags(&m->xs[i], hipStreamNonBlocking);
    for (int i = 0; i < 4 && e == hipSuccess; i++) e = hipEventCreate(&m->ev[i]);
    for (int i = 0; i < 4 && e == hipSuccess; i++)
        e = hipEventCreateWithFlags(&m->evfj[i], hipEventDisableTiming);
    if (e != hipSuccess) {
        delete m;
        return fail(OTFWI_E_HIP, "stream/event creation failed: %s", hipGetErrorString(e));
    }
    *out = m;
    return OTFWI_OK;
}

int otfwi_destroy(otfwi_model *m)
{
    if (!m) return OTFWI_OK;
    // Order: no work in flight on any stream; then the instantiated graphs
    // (they were captured with the fork / join events); then the events and
    // streams; then the memory.
    (void)hipStreamSynchronize(m->stream);
    for (int i = 0; i < 3; i++)
        if (m->xs[i]) (void)hipStreamSynchronize(m->xs[i]);
    // every stream must be idle before anything it may still read is
    // released; if one is not, keep the model's memory (a leak, never a
    // use-after-free) and report it
    bool idle = hipStreamQuery(m->stream) == hipSuccess;
    for (int i = 0; i < 3; i++)
        if (m->xs[i]) idle = idle && hipStreamQuery(m->xs[i]) == hipSuccess;
    if (!idle) {
        fprintf(stderr, "otfwi_destroy: a stream of the model is not idle; its memory is kept\n");
        return fail(OTFWI_E_STATE, "destroy: a stream of the model is not idle");
    }
    for (int i = 0; i < 3; i++) {
        if (m->gexec[i]) (void)hipGraphExecDestroy(m->gexec[i]);
        for (int k = 0; k < 4; k++)
            if (m->gsplit[i][k]) (void)hipGraphExecDestroy(m->gsplit[i][k]);
    }
    for (int i = 0; i < 4; i++)
        if (m->ev[i]) (void)hipEventDestroy(m->ev[i]);
    for (int i = 0; i < 4; i++)
        if (m->evfj[i]) (void)hipEventDestroy(m->evfj[i]);
    for (int i = 0; i < 3; i++)
        if (m->xs[i]) (void)hipStreamDestroy(m->xs[i]);
    (void)hipStreamDestroy(m->stream);
    free_batch(m->b);
    dfree(m->cm); dfree(m->kx); dfree(m->kz); dfree(m->icum); dfree(m->cq32);
    dfree(m->evt); dfree(m->pts); dfree(m->tf_fwd); dfree(m->tf_adj); dfree(m->evrow);
    dfree(m->pmap); dfree(m->ringtab); dfree(m->hilb);
    dfree(m->gacc); dfree(m->gout); dfree(m->ill);
    dfree(m->xsnap); dfree(m->xseg); dfree(m->rxmap); dfree(m->dflag);
    delete m;
    return OTFWI_OK;
}

int otfwi_set_damping(otfwi_model *m, const float *d_x, const float *d_z)
{
    RCK(check_model(m));
    if (!d_x || !d_z) return fail(OTFWI_E_ARG, "damping arrays are NULL");
    const int X = m->X, Z = m->Z;
    for (int z = 1; z < Z; z++)
        for (int x = 0; x < X; x++)
            if (d_x[(size_t)z * X + x] != d_x[x])
                return fail(OTFWI_E_ARG, "d_x varies along z at (x=%d, z=%d): only separable "
                                         "profiles are supported", x, z);
    for (int z = 0; z < Z; z++)
        for (int x = 1; x < X; x++)
            if (d_z[(size_t)z * X + x] != d_z[(size_t)z * X])
                return fail(OTFWI_E_ARG, "d_z varies along x at (x=%d, z=%d): only separable "
                                         "profiles are supported", x, z);
    for (int x = 0; x < X; x++) {
        const float v = x + 1 < X ? d_x[x + 1] : 0.0f;
        if (!std::isfinite(v) || v < 0.0f) return fail(OTFWI_E_NUMERIC, "bad d_x value %g", v);
        m->dxs_h[x] = v;
    }
    for (int z = 0; z < Z; z++) {
        const float v = z + 1 < Z ? d_z[(size_t)(z + 1) * X] : 0.0f;
        if (!std::isfinite(v) || v < 0.0f) return fail(OTFWI_E_NUMERIC, "bad d_z value %g", v);
        m->dzs_h[z] = v;
    }
    RCK(m->prec == 8 ? upload_damping<double>(m) : upload_damping<float>(m));
    m->have_damp = true;
    return OTFWI_OK;
}

int otfwi_set_cquad(otfwi_model *m, const float *cquad)
{
    RCK(check_model(m));
    return set_cq_host(m, cquad);
}

int otfwi_set_cquad_device(otfwi_model *m, const float *cquad_dev)
{
    RCK(check_model(m));
    if (!cquad_dev) return fail(OTFWI_E_ARG, "cquad is NULL");
    const size_t n = (size_t)m->X * m->Z;
    if (!m->cq32) RCK(dalloc(&m->cq32, n));
    if (m->prec == 4) {
        // FP32 build: the model array is formed on the device, nothing goes
        // through host memory but the validity flag
        const size_t np = (size_t)m->zp * m->ld;
        if (!m->cm) {
            RCK(dalloc((char **)&m->cm, np * sizeof(float)));
            HIPCK(hipMemsetAsync(m->cm, 0, np * sizeof(float), m->stream));
        }
        if (!m->dflag) RCK(dalloc(&m->dflag, 1));
        HIPCK(hipMemcpyAsync(m->cq32, cquad_dev, n * sizeof(float), hipMemcpyDeviceToDevice,
                             m->stream));
        // no bad cell: the flag keeps 0x7F7F7F7F (above any 1 + flat index)
        HIPCK(hipMemsetAsync(m->dflag, 0x7F, sizeof(int), m->stream));
        HIPCK(launch_model_cm<float>(m->cq32, m->X, m->Z, m->ld, m->dt, m->dx, (float *)m->cm,
                                     m->dflag, m->stream));
        int bad = 0;
        HIPCK(hipMemcpyAsync(&bad, m->dflag, sizeof(int), hipMemcpyDeviceToHost, m->stream));
        HIPCK(hipStreamSynchronize(m->stream));
        m->cq64_h.clear();
        if (bad != 0x7F7F7F7F) {
            float v = 0.0f;
            HIPCK(hipMemcpy(&v, m->cq32 + (bad - 1), sizeof(float), hipMemcpyDeviceToHost));
            m->have_cq = false;
            return fail(OTFWI_E_NUMERIC, "c^2 at flat index %d is %g (must be finite and >= 0)",
                        bad - 1, (double)v);
        }
        m->have_cq = true;
        return OTFWI_OK;
    }
    HIPCK(hipMemcpy(m->cq32, cquad_dev, n * sizeof(float), hipMemcpyDeviceToDevice));
    std::vector<float> h(n);
    HIPCK(hipMemcpy(h.data(), m->cq32, n * sizeof(float), hipMemcpyDeviceToHost));
    RCK(check_cq(h.data(), n));
    m->have_cq = true;
    m->cq64_h.clear();
    return m->prec == 8 ? upload_model_arrays<double>(m) : upload_model_arrays<float>(m);
}

int otfwi_set_sources(otfwi_model *m, int n_src, const int *src_x, const int *src_z,
                      const float *wavelets)
{
    RCK(check_model(m));
    RCK(check_points(m, n_src, src_x, src_z, "source"));
    if (n_src > 0 && !wavelets) return fail(OTFWI_E_ARG, "wavelets are NULL");
    m->n_src = n_src;
    m->sx.assign(src_x, src_x + n_src);
    m->sz.assign(src_z, src_z + n_src);
    m->src_h.assign(wavelets, wavelets + (size_t)n_src * m->T);
    for (float v : m->src_h)
        if (!std::isfinite(v)) return fail(OTFWI_E_NUMERIC, "non-finite source sample");
    m->points_dirty = true;
    return m->prec == 8 ? upload_sources<double>(m) : upload_sources<float>(m);
}

int otfwi_set_receivers(otfwi_model *m, int n_rx, const int *rx_x, const int *rx_z)
{
    RCK(check_model(m));
    RCK(check_points(m, n_rx, rx_x, rx_z, "receiver"));
    m->n_rx = n_rx;
    m->rxx.assign(rx_x, rx_x + n_rx);
    m->rxz.assign(rx_z, rx_z + n_rx);
    m->points_dirty = true;
    free_batch(m->b);
    return OTFWI_OK;
}

int otfwi_set_ring(otfwi_model *m, int n_ring, const int *ring_x, const int *ring_z)
{
    RCK(check_model(m));
    RCK(check_points(m, n_ring, ring_x, ring_z, "ring point"));
    m->n_ring = n_ring;
    m->vx.assign(ring_x, ring_x + n_ring);
    m->vz.assign(ring_z, ring_z + n_ring);
    m->points_dirty = true;
    free_batch(m->b);
    return OTFWI_OK;
}

int otfwi_set_observed(otfwi_model *m, int n_shots, const float *obs)
{
    RCK(check_model(m));
    if (n_shots <= 0 || !obs) return fail(OTFWI_E_ARG, "bad observed data");
    if (m->n_rx <= 0) return fail(OTFWI_E_STATE, "set receivers before observed data");
    const size_t cnt = (size_t)n_shots * m->n_rx * m->T;
    m->obs_h.assign(obs, obs + cnt);
    m->obs_shots = n_shots;
    m->obs_ver++;
    return OTFWI_OK;
}

int otfwi_set_tuning(otfwi_model *m, int th_fwd, int th_adj, int max_batch)
{
    RCK(check_model(m));
    auto ok = [&](int th, bool adj) {
        if (th == 0) return true;
        if (m->prec == 8) return th == 16 || (!adj && th == 32);
        return th == 16 || th == 32 || th == 64;
    };
    if (!ok(th_fwd, false) || !ok(th_adj, true))
        return fail(OTFWI_E_ARG, "unsupported tile height (fwd %d, adj %d) for precision %d",
                    th_fwd, th_adj, m->prec);
    if (max_batch < 0) return fail(OTFWI_E_ARG, "max_batch must be >= 0");
    if (th_fwd) m->th_fwd = th_fwd;
    if (th_adj) m->th_adj = th_adj;
    if (max_batch != m->max_batch) free_batch(m->b);
    m->max_batch = max_batch;
    return OTFWI_OK;
}

int otfwi_set_launch(otfwi_model *m, int persistent_fwd, int persistent_adj)
{
    RCK(check_model(m));
    m->pers_fwd = persistent_fwd ? 1 : 0;
    m->pers_adj = persistent_adj ? 1 : 0;
    return OTFWI_OK;
}

int otfwi_set_blocking(otfwi_model *m, int steps_fwd, int shots_fwd, int steps_adj,
                       int shots_adj)
{
    RCK(check_model(m));
    if (steps_fwd < 1 || steps_adj < 1 || shots_fwd < 0 || shots_adj < 0)
        return fail(OTFWI_E_ARG, "blocking: steps >= 1 and shots >= 0 (0 = all)");
    m->tb_steps[0] = steps_fwd; m->tb_shots[0] = shots_fwd;
    m->tb_steps[1] = steps_adj; m->tb_shots[1] = shots_adj;
    return OTFWI_OK;
}

int otfwi_set_fwd_steps(otfwi_model *m, int steps)
{
    RCK(check_model(m));
    if (steps < 0 || steps > 2)
        return fail(OTFWI_E_ARG, "steps per forward launch must be 0 (by size), 1 or 2");
    m->fwd_steps = steps;
    return OTFWI_OK;
}

int otfwi_get_fwd_steps(otfwi_model *m, int *steps)
{
    RCK(check_model(m));
    if (!steps) return fail(OTFWI_E_ARG, "output is NULL");
    *steps = m->fwd_steps;
    return OTFWI_OK;
}

int otfwi_steps_per_launch(otfwi_model *m, int n_shots, int *fwd, int *adj)
{
    RCK(check_model(m));
    if (n_shots <= 0 || !fwd || !adj) return fail(OTFWI_E_ARG, "steps_per_launch: bad arguments");
    RCK(ready(m, false));
    const long long plane = (long long)m->zp * m->ld;
    const bool ok = pair2_supported(plane, (long long)std::max(m->b.cap, n_shots) * plane,
                                    m->prec);
    *fwd = ok && use_pairs(m, n_shots, false) ? 2 : 1;
    *adj = ok && use_pairs(m, n_shots, true) ? 2 : 1;
    return OTFWI_OK;
}

int otfwi_loop_modes(otfwi_model *m, int n_shots, int *out)
{
    RCK(check_model(m));
    if (n_shots <= 0 || !out) return fail(OTFWI_E_ARG, "loop_modes: bad arguments");
    int f = 1, a = 1;
    RCK(otfwi_steps_per_launch(m, n_shots, &f, &a));
    const int ns = pair_split(m, n_shots);
    const int fs = f == 2 ? ns : 1, as = a == 2 ? ns : 1;
    out[0] = fs;
    out[1] = m->use_graph ? 1 : 0;
    out[2] = as;
    out[3] = m->use_graph ? 1 : 0;
    return OTFWI_OK;
}

int otfwi_set_pair_streams(otfwi_model *m, int k)
{
    RCK(check_model(m));
    if (k != 0 && k != 1 && k != 2 && k != 4)
        return fail(OTFWI_E_ARG, "pair streams must be 0 (by size), 1, 2 or 4");
    m->pair_streams = k;
    return OTFWI_OK;
}

int otfwi_set_pair_width(otfwi_model *m, int nw)
{
    RCK(check_model(m));
    if (nw != 0 && f2_own(nw) <= 0)
        return fail(OTFWI_E_ARG, "two-step block width must be 0 (automatic), 128, 192, 256 or 384");
    if (nw != m->pair_nw) {
        m->pair_nw = nw;
        m->points_dirty = true;   // tables are rebuilt for the new width
    }
    return OTFWI_OK;
}

int otfwi_set_adj_steps(otfwi_model *m, int steps)
{
    RCK(check_model(m));
    if (steps < 0 || steps > 2)
        return fail(OTFWI_E_ARG, "steps per adjoint launch must be 0 (by size), 1 or 2");
    m->adj_steps = steps;
    return OTFWI_OK;
}

int otfwi_get_adj_steps(otfwi_model *m, int *steps)
{
    RCK(check_model(m));
    if (!steps) return fail(OTFWI_E_ARG, "output is NULL");
    *steps = m->adj_steps;
    return OTFWI_OK;
}

int otfwi_get_blocking(otfwi_model *m, int *out)
{
    RCK(check_model(m));
    if (!out) return fail(OTFWI_E_ARG, "output is NULL");
    out[0] = m->tb_steps[0]; out[1] = m->tb_shots[0];
    out[2] = m->tb_steps[1]; out[3] = m->tb_shots[1];
    return OTFWI_OK;
}

int otfwi_get_launch(otfwi_model *m, int *persistent_fwd, int *persistent_adj)
{
    RCK(check_model(m));
    if (persistent_fwd) *persistent_fwd = m->pers_fwd;
    if (persistent_adj) *persistent_adj = m->pers_adj;
    return OTFWI_OK;
}

int otfwi_set_graph(otfwi_model *m, int use_graph)
{
    RCK(check_model(m));
    m->use_graph = use_graph != 0;
    return OTFWI_OK;
}

int otfwi_autotune(otfwi_model *m, int n_shots, int steps, int *choice, float *ms)
{
    RCK(check_model(m));
    return DISPATCH(m, autotune_impl)(m, n_shots, steps, choice, ms);
}

int otfwi_kernel_times(otfwi_model *m, int n_shots, int reps, float *ms)
{
    RCK(check_model(m));
    if (!ms) return fail(OTFWI_E_ARG, "ms is NULL");
    return DISPATCH(m, kernel_times_impl)(m, n_shots, reps, ms);
}

int otfwi_get_tuning(otfwi_model *m, int *th_fwd, int *th_adj, int *max_batch)
{
    RCK(check_model(m));
    if (th_fwd) *th_fwd = m->th_fwd;
    if (th_adj) *th_adj = m->th_adj;
    if (max_batch) *max_batch = m->max_batch;
    return OTFWI_OK;
}

int otfwi_get_timing(otfwi_model *m, double *out)
{
    RCK(check_model(m));
    if (!out) return fail(OTFWI_E_ARG, "out is NULL");
    out[0] = m->t_fwd_ms;
    out[1] = m->t_mis_ms;
    out[2] = m->t_adj_ms;
    out[3] = (double)m->n_fwd_launch;
    out[4] = (double)m->n_adj_launch;
    return OTFWI_OK;
}

int otfwi_forward(otfwi_model *m, int n_shots, const int *shot_src, float *rec_out)
{
    RCK(check_model(m));
    if (n_shots <= 0) return fail(OTFWI_E_ARG, "n_shots must be positive");
    return DISPATCH(m, forward_impl)(m, n_shots, shot_src, rec_out);
}

int otfwi_gradient(otfwi_model *m, int n_shots, const int *shot_src, int misfit_kind,
                   double *misfit_out, float *grad_out, float *rec_out)
{
    RCK(check_model(m));
    return DISPATCH(m, gradient_impl)(m, n_shots, shot_src, misfit_kind, misfit_out, grad_out,
                                      false, rec_out);
}

int otfwi_gradient_device(otfwi_model *m, int n_shots, const int *shot_src, int misfit_kind,
                          double *misfit_out, float *grad_out_dev)
{
    RCK(check_model(m));
    return DISPATCH(m, gradient_impl)(m, n_shots, shot_src, misfit_kind, misfit_out, grad_out_dev,
                                      true, nullptr);
}

int otfwi_gradient_multi(int n_models, otfwi_model *const *models, const int *n_shots,
                         const int *const *shot_src, const int *misfit_kind,
                         double *const *misfit_out, float *const *grad_out, int n_workers)
{
    if (n_models <= 0 || !models || !n_shots || !misfit_kind)
        return fail(OTFWI_E_ARG, "gradient_multi: bad arguments");
    for (int i = 0; i < n_models; i++) {
        RCK(check_model(models[i]));
        for (int j = 0; j < i; j++)
            if (models[j] == models[i])
                return fail(OTFWI_E_ARG, "gradient_multi: model %d is listed twice", i);
    }
    // longest processing time first: cost = cells x steps x shots
    std::vector<int> order(n_models);
    for (int i = 0; i < n_models; i++) order[i] = i;
    auto cost = [&](int i) {
        const otfwi_model *m = models[i];
        return (double)m->X * m->Z * m->T * std::max(n_shots[i], 1);
    };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost(a) > cost(b); });
    const int nw = std::max(1, std::min(n_workers > 0 ? n_workers : 4, n_models));
    std::atomic<int> next{0};
    std::vector<int> rc(n_models, OTFWI_OK);
    std::vector<std::string> msg(n_models);
    auto worker = [&]() {
        for (;;) {
            const int q = next.fetch_add(1);
            if (q >= n_models) return;
            const int i = order[q];
            otfwi_model *m = models[i];
            rc[i] = DISPATCH(m, gradient_impl)(m, n_shots[i], shot_src ? shot_src[i] : nullptr,
                                               misfit_kind[i], misfit_out ? misfit_out[i] : nullptr,
                                               grad_out ? grad_out[i] : nullptr, false, nullptr);
            if (rc[i] != OTFWI_OK) msg[i] = g_err;
        }
    };
    std::vector<std::thread> pool;
    for (int w = 1; w < nw; w++) pool.emplace_back(worker);
    worker();
    for (auto &t : pool) t.join();
    for (int i = 0; i < n_models; i++)
        if (rc[i] != OTFWI_OK) return fail(rc[i], "model %d: %s", i, msg[i].c_str());
    return OTFWI_OK;
}

int otfwi_gradient_fused(int n_models, otfwi_model *const *models, const int *n_shots,
                         const int *const *shot_src, const int *misfit_kind,
                         double *const *misfit_out, float *const *grad_out)
{
    if (n_models <= 0 || !models || !n_shots || !misfit_kind)
        return fail(OTFWI_E_ARG, "gradient_fused: bad arguments");
    for (int i = 0; i < n_models; i++) {
        RCK(check_model(models[i]));
        for (int j = 0; j < i; j++)
            if (models[j] == models[i])
                return fail(OTFWI_E_ARG, "gradient_fused: model %d is listed twice", i);
    }
    return fused_gradient_impl(n_models, models, n_shots, shot_src, misfit_kind, misfit_out,
                               grad_out);
}

int otfwi_forward_for_grad(otfwi_model *m, int n_shots, const int *shot_src, float *rec_out)
{
    RCK(check_model(m));
    return DISPATCH(m, forward_for_grad_impl)(m, n_shots, shot_src, rec_out);
}

int otfwi_adjoint_from_host(otfwi_model *m, const float *adj_src, float *grad_out)
{
    RCK(check_model(m));
    return DISPATCH(m, adjoint_from_host_impl)(m, adj_src, grad_out);
}

int otfwi_objective(otfwi_model *m, int n_shots, const int *shot_src, int misfit_kind,
                    double *misfit_out)
{
    RCK(check_model(m));
    if (!misfit_out) return fail(OTFWI_E_ARG, "misfit_out is NULL");
    return DISPATCH(m, objective_impl)(m, n_shots, shot_src, misfit_kind, misfit_out);
}

int otfwi_gradient_exact(otfwi_model *m, int n_shots, const int *shot_src, int misfit_kind,
                         int snap_every, double *misfit_out, float *grad_out, double *grad64_out)
{
    RCK(check_model(m));
    return DISPATCH(m, exact_gradient_impl)(m, n_shots, shot_src, misfit_kind, snap_every,
                                            misfit_out, grad_out, grad64_out, false);
}

int otfwi_gradient_exact_device(otfwi_model *m, int n_shots, const int *shot_src, int misfit_kind,
                                int snap_every, double *misfit_out, float *grad_out_dev)
{
    RCK(check_model(m));
    if (!grad_out_dev) return fail(OTFWI_E_ARG, "grad_out_dev is NULL");
    return DISPATCH(m, exact_gradient_impl)(m, n_shots, shot_src, misfit_kind, snap_every,
                                            misfit_out, grad_out_dev, nullptr, true);
}

int otfwi_set_cquad64(otfwi_model *m, const double *cquad)
{
    RCK(check_model(m));
    return set_cq64_host(m, cquad);
}

int otfwi_set_illum(otfwi_model *m, int stride)
{
    RCK(check_model(m));
    if (stride < 0 || stride > 1000000) return fail(OTFWI_E_ARG, "illumination stride %d", stride);
    m->ill_stride = stride;
    m->ill_valid = false;
    return OTFWI_OK;
}

int otfwi_get_illum(otfwi_model *m, float *out, int out_on_device)
{
    RCK(check_model(m));
    if (!out) return fail(OTFWI_E_ARG, "output is NULL");
    if (!m->ill_valid)
        return fail(OTFWI_E_STATE, "no illumination collected (set a stride, then run a gradient)");
    const size_t bytes = (size_t)m->X * m->Z * sizeof(float);
    HIPCK(hipMemcpyAsync(out, m->ill, bytes,
                         out_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                         m->stream));
    HIPCK(hipStreamSynchronize(m->stream));
    return OTFWI_OK;
}

namespace {
int misfit_host(int kind, bool exact, int n_traces, int T, const double *sim, const double *obs,
                double *mis_out, double *adj_out)
{
    if (!misfit_kind_ok(kind) || (exact && kind != OTFWI_MISFIT_L2 && kind != OTFWI_MISFIT_W2))
        return fail(OTFWI_E_ARG, "unknown misfit kind %d%s", kind,
                    exact ? " (the exact derivative supports L2 and W2)" : "");
    if (n_traces <= 0 || T < 2 || !sim || !obs)
        return fail(OTFWI_E_ARG, "bad misfit arguments (n_traces=%d, T=%d)", n_traces, T);
    if (w2_family(kind) && (size_t)2 * T * sizeof(double) > 160 * 1024)
        return fail(OTFWI_E_ARG, "W2 supports T <= %d", (int)(160 * 1024 / 16));
    const size_t cnt = (size_t)n_traces * T;
    double *dsim = nullptr, *dobs = nullptr, *dadj = nullptr, *dmis = nullptr, *dscr = nullptr;
    double *dtab = nullptr;
    int rc = OTFWI_OK;
    auto cleanup = [&] {
        dfree(dsim); dfree(dobs); dfree(dadj); dfree(dmis); dfree(dscr); dfree(dtab);
    };
    if ((rc = dalloc(&dsim, cnt)) || (rc = dalloc(&dobs, cnt)) || (rc = dalloc(&dadj, cnt)) ||
        (rc = dalloc(&dmis, (size_t)n_traces)) || (rc = dalloc(&dscr, cnt * misfit_scratch_per_sample(kind) + n_traces)) ||
        (kind == OTFWI_MISFIT_ENV && (rc = dalloc(&dtab, (size_t)T)))) {
        cleanup();
        return rc;
    }
    int *flags = misfit_flags(kind, dscr, n_traces, T);
    hipError_t e = hipMemcpy(dsim, sim, cnt * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dobs, obs, cnt * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemset(dscr, 0, (cnt * misfit_scratch_per_sample(kind) + n_traces) * sizeof(double));
    if (e == hipSuccess && dtab) {
        std::vector<double> tab(T);
        hilbert_table(T, tab.data());
        e = hipMemcpy(dtab, tab.data(), T * sizeof(double), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess)
        e = exact ? launch_misfit_exact<double, double>(kind, dsim, dobs, dadj, dmis, n_traces, T,
                                                        dscr, flags, nullptr)
                  : launch_misfit<double, double>(kind, dsim, dobs, dadj, nullptr, dmis, n_traces,
                                                  T, dscr, nullptr, dtab);
    std::vector<int> fl(n_traces, 0);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess && mis_out)
        e = hipMemcpy(mis_out, dmis, n_traces * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess && adj_out)
        e = hipMemcpy(adj_out, dadj, cnt * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess && w2_family(kind))
        e = hipMemcpy(fl.data(), flags, n_traces * sizeof(int), hipMemcpyDeviceToHost);
    cleanup();
    if (e != hipSuccess) return fail(OTFWI_E_HIP, "misfit: %s", hipGetErrorString(e));
    // an all-zero observed trace (flag 2) yields NaN, as the reference does
    for (int i = 0; i < n_traces; i++) {
        if (fl[i] == 1) return fail(OTFWI_E_NUMERIC, "W2: observed trace %d has a non-finite CDF", i);
        if (fl[i] == 3)
            return fail(OTFWI_E_NUMERIC, "exact W2: CDF of simulated trace %d is not monotone", i);
    }
    return OTFWI_OK;
}
}  // namespace

int otfwi_misfit(int kind, int n_traces, int T, const double *sim, const double *obs,
                 double *mis_out, double *adj_out)
{
    return misfit_host(kind, false, n_traces, T, sim, obs, mis_out, adj_out);
}

int otfwi_misfit_exact(int kind, int n_traces, int T, const double *sim, const double *obs,
                       double *mis_out, double *rho_out)
{
    return misfit_host(kind, true, n_traces, T, sim, obs, mis_out, rho_out);
}

// ---------------------------------------------------------------------------
// Allocation-free misfit on device arrays: the W2 kernels' per-trace work
// arrays live in a fixed file-scope device array (one copy per GPU, held
// until the process ends); its address on each device is taken once with
// hipGetSymbolAddress and only that address is handed to kernels and copies.
// ---------------------------------------------------------------------------
}  // extern "C"

namespace {

constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)32 << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;

__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
void *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// Address of this device's copy of `scratch` (looked up once per device,
// with that device current).
int scratch_on_current_device(void **out)
{
    int dev = -1;
    HIPCK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return fail(OTFWI_E_ARG, "device %d beyond the %d the library tracks", dev,
                    NOALLOC_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(OTFWI_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                        hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return OTFWI_OK;
}

// Fill n ints with v (the clears a captured graph may replay; no memset nodes).
__global__ void fill_int_kernel(int *p, int n, int v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// status = min(status, base + i) over the traces i whose W2 flag is 1 (an
// observed trace with a non-finite CDF).
__global__ void w2_status_kernel(const int *flags, int n, int base, int *status)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flags[i] == 1) atomicMin(status, base + i);
}

}  // namespace

extern "C" {

int otfwi_misfit_noalloc(int kind, int n_traces, int T, const double *sim, const double *obs,
                         double *mis_out, double *adj_out, void *stream)
{
    if (kind != OTFWI_MISFIT_L2 && kind != OTFWI_MISFIT_W2)
        return fail(OTFWI_E_ARG,
                    "misfit_noalloc: kind %d not supported (L2 = 0 and W2 = 1; the envelope "
                    "misfit needs a Hilbert table: use otfwi_misfit)", kind);
    if (n_traces <= 0 || T < 2 || !sim || !obs || !mis_out || !adj_out)
        return fail(OTFWI_E_ARG, "misfit_noalloc: bad arguments (n_traces=%d, T=%d)", n_traces, T);
    const size_t per_trace = (size_t)3 * T * sizeof(double) + sizeof(int);
    const size_t avail = NOALLOC_SCRATCH_BYTES - 256;
    if (kind == OTFWI_MISFIT_W2 &&
        ((size_t)2 * T * sizeof(double) > 160 * 1024 || per_trace > avail))
        return fail(OTFWI_E_ARG, "misfit_noalloc: W2 supports T <= %d (T = %d)",
                    (int)(160 * 1024 / 16), T);
    hipStream_t st = (hipStream_t)stream;
    if (kind == OTFWI_MISFIT_L2) {
        HIPCK((launch_misfit<double, double>(kind, sim, obs, adj_out, nullptr, mis_out, n_traces,
                                             T, nullptr, st)));
        HIPCK(hipStreamSynchronize(st));
        return OTFWI_OK;
    }
    void *base = nullptr;
    RCK(scratch_on_current_device(&base));
    // scratch: [status int, padded to 256 B][3 nc T doubles][nc flag ints] per chunk
    int *status = (int *)base;
    double *work = (double *)((unsigned char *)base + 256);
    const int chunk = (int)std::min<size_t>(avail / per_trace, (size_t)n_traces);
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, st, status, 1, INT_MAX);
    HIPCK(hipGetLastError());
    for (int c0 = 0; c0 < n_traces; c0 += chunk) {
        const int nc = std::min(chunk, n_traces - c0);
        int *flags = (int *)(work + (size_t)3 * nc * T);
        hipLaunchKernelGGL(fill_int_kernel, dim3((nc + 255) / 256), dim3(256), 0, st, flags, nc, 0);
        HIPCK(hipGetLastError());
        const size_t off = (size_t)c0 * T;
        HIPCK((launch_misfit<double, double>(kind, sim + off, obs + off, adj_out + off, nullptr,
                                             mis_out + c0, nc, T, work, st)));
        hipLaunchKernelGGL(w2_status_kernel, dim3((nc + 255) / 256), dim3(256), 0, st, flags, nc,
                           c0, status);
        HIPCK(hipGetLastError());
    }
    int first_bad = INT_MAX;
    HIPCK(hipMemcpyAsync(&first_bad, status, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCK(hipStreamSynchronize(st));
    if (first_bad != INT_MAX)
        return fail(OTFWI_E_NUMERIC, "W2: observed trace %d has a non-finite CDF", first_bad);
    return OTFWI_OK;
}

// ===========================================================================
// Reference ABI (regression_PML_single.cu), on one process-global FP32 model.
// ===========================================================================

static otfwi_model *g_ref = nullptr;
static float *g_rec_h = nullptr, *g_grad_h = nullptr, *g_sim_h = nullptr, *g_adj_h = nullptr;
static void *g_py_adj = nullptr;

static void ref_report(int rc, const char *where)
{
    if (rc != OTFWI_OK) fprintf(stderr, "otfwi: %s: %s\n", where, g_err.c_str());
}

void init_memory_sim(int X, int Z, int T, float *cquad, float *const_vec, float *d_x, float *d_z,
                     int n_src, int *src_x, int *src_z, int n_rx, int *rx_x, int *rx_z, float *src,
                     float *init, float **rec)
{
    (void)init;
    if (g_ref) otfwi_destroy(g_ref);
    g_ref = nullptr;
    if (g_rec_h) (void)hipHostFree(g_rec_h);
    g_rec_h = nullptr;
    int rc = otfwi_create(&g_ref, X, Z, T, 4, const_vec);
    if (!rc) rc = otfwi_set_damping(g_ref, d_x, d_z);
    if (!rc) rc = otfwi_set_cquad(g_ref, cquad);
    if (!rc) rc = otfwi_set_sources(g_ref, n_src, src_x, src_z, src);
    if (!rc) rc = otfwi_set_receivers(g_ref, n_rx, rx_x, rx_z);
    if (!rc && hipHostMalloc((void **)&g_rec_h, (size_t)std::max(n_rx, 1) * T * sizeof(float)) !=
                   hipSuccess)
        rc = fail(OTFWI_E_MEM, "pinned recording buffer");
    if (!rc && g_rec_h) memset(g_rec_h, 0, (size_t)std::max(n_rx, 1) * T * sizeof(float));
    ref_report(rc, "init_memory_sim");
    if (rec) *rec = g_rec_h;
}

void init_memory_reg(float *obs, float **grad, int *rev_x, int *rev_z, int n_rev)
{
    (void)obs;
    int rc = g_ref ? OTFWI_OK : fail(OTFWI_E_STATE, "init_memory_sim was not called");
    if (!rc) rc = otfwi_set_ring(g_ref, n_rev, rev_x, rev_z);
    if (!rc) {
        if (g_grad_h) (void)hipHostFree(g_grad_h);
        if (g_sim_h) (void)hipHostFree(g_sim_h);
        if (g_adj_h) (void)hipHostFree(g_adj_h);
        g_grad_h = g_sim_h = g_adj_h = nullptr;
        const size_t tr = (size_t)std::max(g_ref->n_rx, 1) * g_ref->T;
        if (hipHostMalloc((void **)&g_grad_h, (size_t)g_ref->X * g_ref->Z * sizeof(float)) !=
                hipSuccess ||
            hipHostMalloc((void **)&g_sim_h, tr * sizeof(float)) != hipSuccess ||
            hipHostMalloc((void **)&g_adj_h, tr * sizeof(float)) != hipSuccess)
            rc = fail(OTFWI_E_MEM, "pinned gradient buffers");
        else
            memset(g_grad_h, 0, (size_t)g_ref->X * g_ref->Z * sizeof(float));
    }
    ref_report(rc, "init_memory_reg");
    if (grad) *grad = g_grad_h;
}

void setCquad(float *cquad)
{
    int rc = g_ref ? otfwi_set_cquad(g_ref, cquad)
                   : fail(OTFWI_E_STATE, "init_memory_sim was not called");
    ref_report(rc, "setCquad");
}

void cuda_simulate(int en_out, int idx_source)
{
    (void)en_out;
    int rc = g_ref ? OTFWI_OK : fail(OTFWI_E_STATE, "init_memory_sim was not called");
    if (!rc) rc = otfwi_forward(g_ref, 1, &idx_source, g_rec_h);
    ref_report(rc, "cuda_simulate");
}

void cuda_grad_ext(float *mse, otfwi_adj_fn adj, int idx_source)
{
    int rc = (g_ref && g_grad_h) ? OTFWI_OK
                                 : fail(OTFWI_E_STATE, "init_memory_sim/init_memory_reg not called");
    if (!rc) rc = otfwi_forward_for_grad(g_ref, 1, &idx_source, g_sim_h);
    if (!rc) {
        const size_t tr = (size_t)g_ref->n_rx * g_ref->T;
        memset(g_adj_h, 0, tr * sizeof(float));
        if (adj) adj(g_sim_h, g_adj_h);
        rc = otfwi_adjoint_from_host(g_ref, g_adj_h, g_grad_h);
    }
    if (mse) *mse = 0.0f;
    ref_report(rc, "cuda_grad_ext");
}

void def_python_adj(void *fn) { g_py_adj = fn; }

}  // extern "C"
