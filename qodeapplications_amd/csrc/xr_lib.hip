// Single translation unit of libxrhip.so: the device-side binomial table
// (constant memory) is shared by every kernel without relocatable device code.
#include "xr_ctx.hip"
#include "xr_excite.hip"
#include "xr_gemm.hip"
#include "xr_ozaki.hip"
#include "xr_density.hip"
#include "xr_contract.hip"
#include "xr_pairs.hip"
#include "xr_solver.hip"
#include "xr_spin.hip"
#include "xr_bound.hip"
