// Ozaki-II split-precision DGEMM over INT8 MFMA (placeholder until the kernels land).
#include "common.hpp"

namespace rstsr {
namespace hip {

struct GemmProblem;
void ozaki_gemm_f64(rstsr_hip_ctx* ctx, const GemmProblem& P, double alpha, double beta, bool beta_zero) {
    (void)ctx;
    (void)P;
    (void)alpha;
    (void)beta;
    (void)beta_zero;
    RSTSR_RAISE(UnImplemented, "Ozaki split GEMM not built yet");
}

}  // namespace hip
}  // namespace rstsr
