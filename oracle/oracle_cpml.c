/* CPU oracle of the second-order C-PML pressure scheme
 * (simulator/src/regression_CPML.cu, prec_deriv = 4).  TEST INFRASTRUCTURE
 * ONLY: tests/ load it as the checker; the product library never does.
 *
 * One shot, plain [slot][z][x] planes as the reference keeps them:
 *   P: 4 slots (t % 4), the first two from `initial` (cuda_simulate2, :442-443);
 *   memory variables: the value of step t - 1 is all a step reads, so two
 *   slots stand for the reference's four.
 * Per step t >= 2 (cuda_simulate2, :459-482):
 *   frame 1 (simulateFrame1, :46-120) on 4 <= x < X-4, 4 <= z < Z-4, reading
 *     P(t-1) as 0 outside that box:
 *     vx0 = (float)D+x(P) / dx; mx(t) = bxh mx(t-1) + axh vx0;
 *     vx1 = (vx0 / Kxh + mx(t)) / rho_half_x; likewise z;
 *   frame 2 (simulateFrame2, :123-210) on the same box, reading vx1, vz1 as 0
 *     outside it: vxx0 = (float)D-x(vx1) / dx; mxx(t) = bx mxx(t-1) + ax vxx0;
 *     vxx1 = vxx0 / Kx + mxx(t); likewise z;
 *     P(t) = 2 P(t-1) - P(t-2) + (vxx1 + vzz1) dt2 kappa;
 *   sources (somaFonte, :213-237): P(t)[src n] += qt_pi cquad[src] s[n][t] dt2
 *     for every n when ppe != 1 or idx == -1, else for n == idx only;
 * and for every t (gravaBufferSensores2ordem, :240-248): rec[n][t] = P(t)[rx n].
 * D+ / D- are the 8-point staggered differences of deriv_macros.h:85-108
 * (offset 0 / 1), each term a double literal times a float, summed in double
 * left to right, then rounded to float.  Built with -ffp-contract=off. */
#include <stdlib.h>
#include <string.h>

static const double C1 = 1.196289e+00, C2 = 7.975260e-02, C3 = 9.570313e-03, C4 = 6.975446e-04;

/* derivPMLdx(m, off, x, z) of deriv_macros.h:85-95 on a row */
static float dpml(const float *r, int x, int off)
{
    return (float)(-C4 * r[x + 4 - off] + C3 * r[x + 3 - off] - C2 * r[x + 2 - off] +
                   C1 * r[x + 1 - off] - C1 * r[x + 0 - off] + C2 * r[x - 1 - off] -
                   C3 * r[x - 2 - off] + C4 * r[x - 3 - off]);
}
/* derivPMLdz: the same along z (stride X) */
static float dpmlz(const float *c, int X, int off)
{
    return (float)(-C4 * c[(4 - off) * X] + C3 * c[(3 - off) * X] - C2 * c[(2 - off) * X] +
                   C1 * c[(1 - off) * X] - C1 * c[(0 - off) * X] + C2 * c[(-1 - off) * X] -
                   C3 * c[(-2 - off) * X] + C4 * c[(-3 - off) * X]);
}

/* Masked copy: the box 4 <= x < X-4, 4 <= z < Z-4 of src, 0 elsewhere, with
 * a 4-cell zero frame around the plane so stencils need no bounds. */
static void boxed(const float *src, float *dst, int X, int Z)
{
    const int W = X + 8;
    memset(dst, 0, sizeof(float) * (size_t)W * (Z + 8));
    for (int z = 4; z < Z - 4; z++)
        memcpy(dst + (size_t)(z + 4) * W + 4 + 4, src + (size_t)z * X + 4, sizeof(float) * (X - 8));
}

void orc_cpml_forward(int X, int Z, int T, int n_src, int n_rx, int ppe, const int *sx,
                      const int *sz, const int *rx, const int *rz, float dx, float dz, float dt2,
                      float qt_pi, const float *cquad, const float *kappa, const float *rho_hx,
                      const float *rho_hz, const float *K_x, const float *a_x, const float *b_x,
                      const float *K_xh, const float *a_xh, const float *b_xh, const float *K_z,
                      const float *a_z, const float *b_z, const float *K_zh, const float *a_zh,
                      const float *b_zh, const float *src, const float *initial, int idx,
                      float *rec)
{
    const size_t n = (size_t)X * Z;
    const int W = X + 8;
    float *P = calloc(4 * n, sizeof(float));
    float *mx = calloc(2 * n, sizeof(float)), *mz = calloc(2 * n, sizeof(float));
    float *mxx = calloc(2 * n, sizeof(float)), *mzz = calloc(2 * n, sizeof(float));
    float *vx1 = calloc(n, sizeof(float)), *vz1 = calloc(n, sizeof(float));
    float *bp = malloc(sizeof(float) * (size_t)W * (Z + 8));
    float *bq = malloc(sizeof(float) * (size_t)W * (Z + 8));
    if (initial) memcpy(P, initial, 2 * n * sizeof(float));
    for (int t = 0; t < T; t++) {
        float *Pt = P + (size_t)(t % 4) * n;
        if (t > 1) {
            const float *P1 = P + (size_t)((t - 1) % 4) * n, *P2 = P + (size_t)((t - 2) % 4) * n;
            const int c = t & 1, p = c ^ 1;
            boxed(P1, bp, X, Z);
            for (int z = 4; z < Z - 4; z++)
                for (int x = 4; x < X - 4; x++) {
                    const size_t i = (size_t)z * X + x;
                    const float *row = bp + (size_t)(z + 4) * W + 4;
                    const float v0 = dpml(row, x, 0) / dx;
                    mx[c * n + i] = b_xh[x] * mx[p * n + i] + a_xh[x] * v0;
                    vx1[i] = (v0 / K_xh[x] + mx[c * n + i]) / rho_hx[i];
                    const float w0 = dpmlz(row + x, W, 0) / dz;
                    mz[c * n + i] = b_zh[z] * mz[p * n + i] + a_zh[z] * w0;
                    vz1[i] = (w0 / K_zh[z] + mz[c * n + i]) / rho_hz[i];
                }
            boxed(vx1, bp, X, Z);
            boxed(vz1, bq, X, Z);
            for (int z = 4; z < Z - 4; z++)
                for (int x = 4; x < X - 4; x++) {
                    const size_t i = (size_t)z * X + x;
                    const float v0 = dpml(bp + (size_t)(z + 4) * W + 4, x, 1) / dx;
                    mxx[c * n + i] = b_x[x] * mxx[p * n + i] + a_x[x] * v0;
                    const float vxx = v0 / K_x[x] + mxx[c * n + i];
                    const float w0 = dpmlz(bq + (size_t)(z + 4) * W + 4 + x, W, 1) / dz;
                    mzz[c * n + i] = b_z[z] * mzz[p * n + i] + a_z[z] * w0;
                    const float vzz = w0 / K_z[z] + mzz[c * n + i];
                    const float lap = vxx + vzz;
                    Pt[i] = 2 * P1[i] - P2[i] + lap * dt2 * kappa[i];
                }
            for (int s = 0; s < n_src; s++) {
                if (ppe == 1 && idx != -1 && idx != s) continue;
                const size_t i = (size_t)sz[s] * X + sx[s];
                Pt[i] += qt_pi * cquad[i] * src[(size_t)s * T + t] * dt2;
            }
        }
        for (int r = 0; r < n_rx; r++) rec[(size_t)r * T + t] = Pt[(size_t)rz[r] * X + rx[r]];
    }
    free(P); free(mx); free(mz); free(mxx); free(mzz); free(vx1); free(vz1); free(bp); free(bq);
}
