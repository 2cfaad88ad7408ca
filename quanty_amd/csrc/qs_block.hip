// Block Lanczos (tensor spectra) and lowest eigenstates.
#include "qs_internal.h"

extern "C" qs_status qs_block_lanczos(qs_ctx *ctx, const qs_hop *h, int k, const double *V0, int nsteps,
                                      double sv_cut, double *A_out, double *B_out, int *kept) {
  return qs_fail(ctx, QS_EUNSUPPORTED, "qs_block_lanczos: not built yet");
}

extern "C" qs_status qs_ground_states(qs_ctx *ctx, const qs_hop *h, int nstates, double tol, int max_krylov,
                                      int max_restarts, uint64_t seed, const double *start, double *E_out,
                                      double *vecs, double *variance_out, int *iterations_out) {
  return qs_fail(ctx, QS_EUNSUPPORTED, "qs_ground_states: not built yet");
}

extern "C" qs_status qs_basis_closure(qs_ctx *ctx, const qs_op *op, const qs_basis *seed, int nres,
                                      const uint8_t *masks, const int *lo, const int *hi, int64_t max_dim,
                                      qs_basis **out) {
  return qs_fail(ctx, QS_EUNSUPPORTED, "qs_basis_closure: not built yet");
}

void qs_heev(int n, const double *A, double *w, double *Z) {}
