// inst.hip -- the kernels of one arithmetic type (TW_T) and dimension (TW_D),
// all element orders, behind plain functions the ABI unit dispatches to.
#include "twoldg_impl.h"

#define TW_CAT2(a, b) a##b
#define TW_CAT(a, b) TW_CAT2(a, b)
#define TW_FN(name) TW_CAT(TW_CAT(name, _), TW_CAT(TW_TAG, TW_D))

int TW_FN(twoldg_i_rhs)(twoldg_problem* p, int src, int dst, double t) {
  return dispatch_p<RhsF, TW_T, TW_D>(p, src, dst, t);
}
int TW_FN(twoldg_i_rk4)(twoldg_problem* p, double t0, double dt, int64_t n) {
  return dispatch_p<Rk4F, TW_T, TW_D>(p, t0, dt, n);
}
int TW_FN(twoldg_i_trunc)(twoldg_problem* p, int slot) { return dispatch_p<TruncF, TW_T, TW_D>(p, slot); }
int TW_FN(twoldg_i_energy)(twoldg_problem* p, int slot, double t, double* out) {
  return dispatch_p<EnergyF, TW_T, TW_D>(p, slot, t, out);
}
int TW_FN(twoldg_i_norms)(twoldg_problem* p, int a, int b, int mask, int pn, double* lp, double* li,
                          double* part, double* red) {
  return dispatch_p<NormsF, TW_T, TW_D>(p, a, b, mask, pn, lp, li, part, red);
}
int TW_FN(twoldg_i_modes)(twoldg_problem* p, int slot, int var, ModeArgs M, double* tau) {
  return dispatch_p<ModesF, TW_T, TW_D>(p, slot, var, M, tau);
}
int TW_FN(twoldg_i_transfer)(twoldg_problem* p, int dslot, const twoldg_problem* src, int sslot,
                             const TransferBlock* map) {
  return dispatch_p<TransferF, TW_T, TW_D>(p, dslot, src, sslot, map);
}
