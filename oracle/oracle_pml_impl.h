/*
 * Scheme body of the CPU oracle, instantiated twice by oracle_pml.c:
 *   SFX=_ref : FP32 storage, staggered derivatives evaluated in FP64 and then
 *              rounded to FP32, every update in FP32 (the reference's own mix);
 *   SFX=_f64 : the same scheme with storage and arithmetic in FP64.
 * Macros expected: SFX (name suffix), ST (storage type), AR (update arithmetic
 * type), DV(expr) (conversion of the FP64 derivative sum to AR).
 *
 * TEST INFRASTRUCTURE: this file is the checker, never the product path.
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)

/* Field access with the reference's tile rule: an operand at a cell outside
 * the interior box 4<=x<X-4, 4<=z<Z-4 reads 0
 * (regression_PML_single.cu:10, 489-498, 573-582, 429-437). */
static inline ST CAT(rd, SFX)(const ST *f, int X, int Z, int x, int z)
{
    return in_box(X, Z, x, z) ? f[(size_t)z * X + x] : (ST)0;
}

/* derivPMLdx / derivPMLdz for prec_deriv==4 (deriv_macros.h:85-108): eight
 * FP64 products summed left to right, the +4 tap first.  off=1 is the
 * backward (D-) operator used on A in the P step, off=0 the forward (D+)
 * operator used on P in the A step. */
static inline double CAT(dpx, SFX)(const ST *f, int X, int Z, int x, int z, int off)
{
#define F_(k) ((double)CAT(rd, SFX)(f, X, Z, x + (k) - off, z))
    return -C4 * F_(4) + C3 * F_(3) - C2 * F_(2) + C1 * F_(1)
           - C1 * F_(0) + C2 * F_(-1) - C3 * F_(-2) + C4 * F_(-3);
#undef F_
}

static inline double CAT(dpz, SFX)(const ST *f, int X, int Z, int x, int z, int off)
{
#define F_(k) ((double)CAT(rd, SFX)(f, X, Z, x, z + (k) - off))
    return -C4 * F_(4) + C3 * F_(3) - C2 * F_(2) + C1 * F_(1)
           - C1 * F_(0) + C2 * F_(-1) - C3 * F_(-2) + C4 * F_(-3);
#undef F_
}

/* One P half-step on every interior cell (simulateFrameP,
 * regression_PML_single.cu:447-528).  Reads slot `prv`, writes slot `cur`.
 * revert=1 is the undamped time-reversed update (:517-518). */
static void CAT(p_step, SFX)(int X, int Z, ST *Px, ST *Pz, const ST *Ax, const ST *Az,
                             const ST *cq, const ST *dxm, const ST *dzm,
                             AR dx, AR dt, int prv, int cur, int revert)
{
    const size_t XZ = (size_t)X * Z;
    const ST *Axp = Ax + prv * XZ, *Azp = Az + prv * XZ;
    const ST *Pxp = Px + prv * XZ, *Pzp = Pz + prv * XZ;
    ST *Pxc = Px + cur * XZ, *Pzc = Pz + cur * XZ;
#pragma omp parallel for schedule(static)
    for (int z = BOX; z < Z - BOX; z++) {
        for (int x = BOX; x < X - BOX; x++) {
            const size_t i = (size_t)z * X + x;
            const AR dA_x = DV(CAT(dpx, SFX)(Axp, X, Z, x, z, 1));
            const AR dA_z = DV(CAT(dpz, SFX)(Azp, X, Z, x, z, 1));
            /* damping is looked up one cell further along its axis (:505-506) */
            const AR dxx = (AR)dxm[i + 1];
            const AR dzz = (AR)dzm[i + X];
            const AR m = (AR)cq[i];
            if (!revert) {
                Pxc[i] = (ST)((AR)Pxp[i] * ((AR)1 - dxx * dt) + m * dt / dx * dA_x);
                Pzc[i] = (ST)((AR)Pzp[i] * ((AR)1 - dzz * dt) + m * dt / dx * dA_z);
            } else {
                Pxc[i] = (ST)((AR)Pxp[i] - m * (dt / dx) * dA_x);
                Pzc[i] = (ST)((AR)Pzp[i] - m * (dt / dx) * dA_z);
            }
        }
    }
}

/* One A half-step (simulateFrameA, regression_PML_single.cu:531-615).  The
 * two FP32-rounded derivatives of Px and Pz are summed before scaling. */
static void CAT(a_step, SFX)(int X, int Z, const ST *Px, const ST *Pz, ST *Ax, ST *Az,
                             const ST *dxm, const ST *dzm, AR dx, AR dt,
                             int prv, int cur, int revert)
{
    const size_t XZ = (size_t)X * Z;
    const ST *Pxc = Px + cur * XZ, *Pzc = Pz + cur * XZ;
    const ST *Axp = Ax + prv * XZ, *Azp = Az + prv * XZ;
    ST *Axc = Ax + cur * XZ, *Azc = Az + cur * XZ;
#pragma omp parallel for schedule(static)
    for (int z = BOX; z < Z - BOX; z++) {
        for (int x = BOX; x < X - BOX; x++) {
            const size_t i = (size_t)z * X + x;
            const AR dPxdx = DV(CAT(dpx, SFX)(Pxc, X, Z, x, z, 0));
            const AR dPxdz = DV(CAT(dpz, SFX)(Pxc, X, Z, x, z, 0));
            const AR dPzdx = DV(CAT(dpx, SFX)(Pzc, X, Z, x, z, 0));
            const AR dPzdz = DV(CAT(dpz, SFX)(Pzc, X, Z, x, z, 0));
            const AR dxx = (AR)dxm[i + 1];
            const AR dzz = (AR)dzm[i + X];
            if (!revert) {
                Axc[i] = (ST)((AR)Axp[i] * ((AR)1 - dxx * dt) + dt / dx * (dPxdx + dPzdx));
                Azc[i] = (ST)((AR)Azp[i] * ((AR)1 - dzz * dt) + dt / dx * (dPxdz + dPzdz));
            } else {
                Axc[i] = (ST)((AR)Axp[i] - dt / dx * (dPxdx + dPzdx));
                Azc[i] = (ST)((AR)Azp[i] - dt / dx * (dPxdz + dPzdz));
            }
        }
    }
}

/* Running-sum injection (somaFonteIntegral, :276-292): integral_n += s[n][t]
 * (s[n][T-1-t] when flip), then half of it into each split field. */
static void CAT(inject, SFX)(int X, int Z, int T, ST *Px, ST *Pz, int cur, int n,
                             const int *px, const int *pz, const ST *s, ST *integ,
                             int t, int flip, int idx)
{
    const size_t XZ = (size_t)X * Z;
    for (int k = 0; k < n; k++) {
        if (idx != -1 && idx != k)
            continue;
        integ[k] = (ST)((AR)integ[k] + (AR)s[(size_t)k * T + (flip ? T - 1 - t : t)]);
        const size_t i = cur * XZ + (size_t)pz[k] * X + px[k];
        Px[i] = (ST)((AR)Px[i] + (AR)integ[k] / (AR)2);
        Pz[i] = (ST)((AR)Pz[i] + (AR)integ[k] / (AR)2);
    }
}

/* Point sampling buf[n*T+t] = Px+Pz (gravaBufferSensores, :295-309). */
static void CAT(sample, SFX)(int X, int Z, int T, const ST *Px, const ST *Pz, int cur,
                             int n, const int *px, const int *pz, ST *buf, int t)
{
    const size_t XZ = (size_t)X * Z;
    for (int k = 0; k < n; k++) {
        const size_t i = cur * XZ + (size_t)pz[k] * X + px[k];
        buf[(size_t)k * T + t] = (ST)((AR)Px[i] + (AR)Pz[i]);
    }
}

/* Forward run, cuda_simulate (:618-702) / phase 1 of cuda_grad_ext
 * (:745-769).  rec (n_rx*T) receives Px+Pz at the receivers for t>=1 and 0
 * at t=0; ring (n_rev*T, may be NULL) the same at the ring points; seed
 * (4*X*Z, may be NULL) receives Px,Pz of slot (T-1)%2 and Ax,Az of slot
 * (T-2)%2, which is what phase 2 starts the reverse run from (:787-790). */
int CAT(orc_forward, SFX)(int X, int Z, int T, const ST *cq, const float *cv,
                          const ST *dxm, const ST *dzm,
                          int ns, const int *sx, const int *sz, const ST *src, int idx,
                          int nr, const int *rx, const int *rz, ST *rec,
                          int nv, const int *vx, const int *vz, ST *ring, ST *seed)
{
    if (X < 2 * BOX + 1 || Z < 2 * BOX + 1 || T < 2)
        return -1;
    const size_t XZ = (size_t)X * Z;
    const AR dx = (AR)cv[0], dt = (AR)cv[2];
    ST *Px = calloc(2 * XZ, sizeof(ST)), *Pz = calloc(2 * XZ, sizeof(ST));
    ST *Ax = calloc(2 * XZ, sizeof(ST)), *Az = calloc(2 * XZ, sizeof(ST));
    ST *integ = calloc(ns > 0 ? ns : 1, sizeof(ST));
    if (!Px || !Pz || !Ax || !Az || !integ)
        return -2;
    if (rec)
        memset(rec, 0, (size_t)nr * T * sizeof(ST));
    if (ring)
        memset(ring, 0, (size_t)nv * T * sizeof(ST));
    for (int t = 1; t < T; t++) {
        const int cur = t % 2, prv = (t - 1) % 2;
        CAT(p_step, SFX)(X, Z, Px, Pz, Ax, Az, cq, dxm, dzm, dx, dt, prv, cur, 0);
        CAT(inject, SFX)(X, Z, T, Px, Pz, cur, ns, sx, sz, src, integ, t, 0, idx);
        if (ring)
            CAT(sample, SFX)(X, Z, T, Px, Pz, cur, nv, vx, vz, ring, t);
        if (rec)
            CAT(sample, SFX)(X, Z, T, Px, Pz, cur, nr, rx, rz, rec, t);
        CAT(a_step, SFX)(X, Z, Px, Pz, Ax, Az, dxm, dzm, dx, dt, prv, cur, 0);
    }
    if (seed) {
        memcpy(seed + 0 * XZ, Px + ((T - 1) % 2) * XZ, XZ * sizeof(ST));
        memcpy(seed + 1 * XZ, Pz + ((T - 1) % 2) * XZ, XZ * sizeof(ST));
        memcpy(seed + 2 * XZ, Ax + ((T - 2) % 2) * XZ, XZ * sizeof(ST));
        memcpy(seed + 3 * XZ, Az + ((T - 2) % 2) * XZ, XZ * sizeof(ST));
    }
    free(Px); free(Pz); free(Ax); free(Az); free(integ);
    return 0;
}

/* 9-point 8th-order second difference in FP32 (deriv_x / deriv_z,
 * deriv_macros.h:110-135), tap order +4 .. -4, then / 5040. */
static inline AR CAT(lap, SFX)(const ST *u, int X, int Z, int x, int z)
{
#define UX(k) ((AR)CAT(rd, SFX)(u, X, Z, x + (k), z))
#define UZ(k) ((AR)CAT(rd, SFX)(u, X, Z, x, z + (k)))
    const AR lx = (-(AR)9 * UX(4) + (AR)128 * UX(3) - (AR)1008 * UX(2) + (AR)8064 * UX(1)
                   - (AR)14350 * UX(0) + (AR)8064 * UX(-1) - (AR)1008 * UX(-2)
                   + (AR)128 * UX(-3) - (AR)9 * UX(-4)) / (AR)5040;
    const AR lz = (-(AR)9 * UZ(4) + (AR)128 * UZ(3) - (AR)1008 * UZ(2) + (AR)8064 * UZ(1)
                   - (AR)14350 * UZ(0) + (AR)8064 * UZ(-1) - (AR)1008 * UZ(-2)
                   + (AR)128 * UZ(-3) - (AR)9 * UZ(-4)) / (AR)5040;
#undef UX
#undef UZ
    return lx + lz;
}

/* Phase 2 of cuda_grad_ext (:783-836): time-reversed reconstruction of the
 * forward field from `seed` with ring restore, adjoint propagation with the
 * running-sum of the flipped adjoint source at the receivers, and the
 * imaging condition grad += (Px_f+Pz_f) * lap8(Px_b+Pz_b) at every t.
 * uhist (T*Z*X, may be NULL) receives the reconstructed U = Px_b+Pz_b and
 * lhist (T*Z*X, may be NULL) the adjoint field Px_f+Pz_f of every t, as the
 * imaging condition reads them. */
int CAT(orc_adjoint, SFX)(int X, int Z, int T, const ST *cq, const float *cv,
                          const ST *dxm, const ST *dzm,
                          int nr, const int *rx, const int *rz, const ST *adj,
                          int nv, const int *vx, const int *vz, const ST *ring,
                          const ST *seed, ST *grad, ST *uhist, ST *lhist)
{
    if (X < 2 * BOX + 1 || Z < 2 * BOX + 1 || T < 2)
        return -1;
    const size_t XZ = (size_t)X * Z;
    const AR dx = (AR)cv[0], dt = (AR)cv[2];
    ST *Pxf = calloc(2 * XZ, sizeof(ST)), *Pzf = calloc(2 * XZ, sizeof(ST));
    ST *Axf = calloc(2 * XZ, sizeof(ST)), *Azf = calloc(2 * XZ, sizeof(ST));
    ST *Pxb = calloc(2 * XZ, sizeof(ST)), *Pzb = calloc(2 * XZ, sizeof(ST));
    ST *Axb = calloc(2 * XZ, sizeof(ST)), *Azb = calloc(2 * XZ, sizeof(ST));
    ST *U = calloc(XZ, sizeof(ST)), *L = calloc(XZ, sizeof(ST));
    ST *integ = calloc(nr > 0 ? nr : 1, sizeof(ST));
    if (!Pxf || !Pzf || !Axf || !Azf || !Pxb || !Pzb || !Axb || !Azb || !U || !L || !integ)
        return -2;
    memcpy(Pxb, seed + 0 * XZ, XZ * sizeof(ST));
    memcpy(Pzb, seed + 1 * XZ, XZ * sizeof(ST));
    memcpy(Axb, seed + 2 * XZ, XZ * sizeof(ST));
    memcpy(Azb, seed + 3 * XZ, XZ * sizeof(ST));
    memset(grad, 0, XZ * sizeof(ST));
    for (int t = 0; t < T; t++) {
        const int cur = t % 2, prv = (t + 1) % 2;
        if (t > 0) {
            CAT(p_step, SFX)(X, Z, Pxf, Pzf, Axf, Azf, cq, dxm, dzm, dx, dt, prv, cur, 0);
            CAT(p_step, SFX)(X, Z, Pxb, Pzb, Axb, Azb, cq, dxm, dzm, dx, dt, prv, cur, 1);
            CAT(inject, SFX)(X, Z, T, Pxf, Pzf, cur, nr, rx, rz, adj, integ, t, 1, -1);
            /* ring restore (setRec, :260-273) */
            for (int k = 0; k < nv; k++) {
                const size_t i = cur * XZ + (size_t)vz[k] * X + vx[k];
                const ST v = (ST)((AR)ring[(size_t)k * T + T - 1 - t] / (AR)2);
                Pxb[i] = v;
                Pzb[i] = v;
            }
            CAT(a_step, SFX)(X, Z, Pxf, Pzf, Axf, Azf, dxm, dzm, dx, dt, prv, cur, 0);
            CAT(a_step, SFX)(X, Z, Pxb, Pzb, Axb, Azb, dxm, dzm, dx, dt, prv, cur, 1);
        }
        /* somaFrames (:312-327) then calcGrad2 (:390-445) */
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < XZ; i++) {
            U[i] = (ST)((AR)Pxb[cur * XZ + i] + (AR)Pzb[cur * XZ + i]);
            L[i] = (ST)((AR)Pxf[cur * XZ + i] + (AR)Pzf[cur * XZ + i]);
        }
        if (uhist)
            memcpy(uhist + (size_t)t * XZ, U, XZ * sizeof(ST));
        if (lhist)
            memcpy(lhist + (size_t)t * XZ, L, XZ * sizeof(ST));
#pragma omp parallel for schedule(static)
        for (int z = 0; z < Z; z++)
            for (int x = 0; x < X; x++) {
                const size_t i = (size_t)z * X + x;
                grad[i] = (ST)((AR)grad[i] + (AR)L[i] * CAT(lap, SFX)(U, X, Z, x, z));
            }
    }
    free(Pxf); free(Pzf); free(Axf); free(Azf);
    free(Pxb); free(Pzb); free(Axb); free(Azb);
    free(U); free(L); free(integ);
    return 0;
}

#undef CAT
#undef CAT_
