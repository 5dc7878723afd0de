// Goldstein patch pass on MFMA: DFT-by-GEMM with hi/lo split FP16 operands
// and FP32 accumulation.
#include "s1ifg_internal.h"
#include "gold_dev.h"

namespace s1ifg {

int goldstein_mfma(s1ifg_ctx *ctx, const float2 *in, int nrow, int ncol, float alpha, int pi_n,
                   int pj_n, float2 *patches) {
  (void)ctx; (void)in; (void)nrow; (void)ncol; (void)alpha; (void)pi_n; (void)pj_n; (void)patches;
  return set_error(S1IFG_ERANGE, "FP16-split Goldstein path is not built yet");
}

}  // namespace s1ifg
