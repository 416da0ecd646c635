/*
 * CPU oracle for the 2D acoustic FWI hot path: a plain-C restatement of the
 * reference's split-field-PML propagator and boundary-saving adjoint gradient
 * (simulator/src/regression_PML_single.cu:260-846 with the prec_deriv==4
 * branch of simulator/src/deriv_macros.h:84-136).
 *
 * TEST INFRASTRUCTURE ONLY.  tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg call it as the checker; the product (otfwi_amd) never
 * links, loads or calls it.
 *
 * Two instantiations of the same scheme (oracle_pml_impl.h):
 *   orc_forward_ref / orc_adjoint_ref : FP32 state, FP64-evaluated staggered
 *       derivatives rounded to FP32, FP32 updates -- the reference's mix
 *       (the macro literals carry no f suffix, deriv_macros.h:87-107);
 *   orc_forward_f64 / orc_adjoint_f64 : everything in FP64.
 * Built with -ffp-contract=off so each product and sum rounds separately.
 * Rows are split over OpenMP threads; every cell is computed independently,
 * so results do not depend on the thread count.
 */
#include <stdlib.h>
#include <string.h>
#include <omp.h>

/* prec_deriv==4 staggered coefficients exactly as written in
 * deriv_macros.h:87-94 (truncated decimals of 1225/1024, 245/3072, 49/5120,
 * 5/7168). */
#define C1 1.196289e+00
#define C2 7.975260e-02
#define C3 9.570313e-03
#define C4 6.975446e-04

/* 4-cell outer ring that the reference never updates (`inbounds(x,z,4)`,
 * regression_PML_single.cu:10, 16, 508, 595). */
#define BOX 4

static inline int in_box(int X, int Z, int x, int z)
{
    return x >= BOX && z >= BOX && x < X - BOX && z < Z - BOX;
}

#define SFX _ref
#define ST float
#define AR float
#define DV(e) ((float)(e))
#include "oracle_pml_impl.h"
#undef SFX
#undef ST
#undef AR
#undef DV

#define SFX _f64
#define ST double
#define AR double
#define DV(e) (e)
#include "oracle_pml_impl.h"
#undef SFX
#undef ST
#undef AR
#undef DV

int orc_num_threads(void) { return omp_get_max_threads(); }
void orc_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }
