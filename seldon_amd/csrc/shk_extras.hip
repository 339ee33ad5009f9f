// shk_extras.hip -- entry points beyond Seldon's API: FP32+FP64 iterative
// refinement, the adjoint SDDMM gradient and the batched CG.  (Filled in by
// the dedicated sources; these definitions report "not available" until the
// corresponding kernels are linked.)
#include "shk_internal.hpp"

using namespace shk;

extern "C" {

int shk_solve_ir(shk_ctx, shk_csr, shk_csr, void *, const void *, shk_precond, shk_ir_params *, int *)
{
    return fail(SHK_ERR_STATE, "shk_solve_ir: not available in this build");
}

int shk_sddmm(shk_ctx, shk_csr, const void *, const void *, void *)
{
    return fail(SHK_ERR_STATE, "shk_sddmm: not available in this build");
}

int shk_cg_batched(shk_ctx, shk_csr, int, const int *, void *, const void *, double, int, int, int, int *,
                   int *)
{
    return fail(SHK_ERR_STATE, "shk_cg_batched: not available in this build");
}

int shk_cg_batched_ir(shk_ctx, shk_csr, shk_csr, int, const int *, void *, const void *, double, int, double,
                      int, int, int *, int *)
{
    return fail(SHK_ERR_STATE, "shk_cg_batched_ir: not available in this build");
}

}  // extern "C"
