// Internal plans and helpers of the density pipeline.
#pragma once
#include "xr_common.h"
#include <cstring>

// One side of a factorised operator string: k operators of one kind
// (is_c = 1 creators, 0 annihilators); slot[j] is the position, in the
// ascending index tuple, of the orbital that operator j (leftmost first) uses.
struct SidePlan {
    int k;
    int is_c;
    int slot[3];
};

// op = L * R.  L is stored through its adjoint (annihilators acting on the
// bra determinant); hasL = 0 means no L part.
struct OpPlan {
    int k;
    int hasL;
    SidePlan L;
    SidePlan R;
    int delta;  // bra electrons - ket electrons
};

int xr_plan_op(const char* ops, OpPlan* p);
int xr_space_masks(xr_ctx* c, int n_orbs, int n_core, int n_elec, const uint64_t** masks,
                   int64_t* n);
int xr_transpose(xr_ctx* c, const double* z, int64_t nS, int64_t nC, double* zT);
int xr_side_table(xr_ctx* c, const SidePlan& sd, const uint64_t* Km, int64_t nK, int n_orbs,
                  int n_core, int32_t** tab);
int xr_gather(xr_ctx* c, const int32_t* tab, int64_t nK, int64_t nT, const double* zT,
              int64_t nS, double* out, int layout);
int xr_index_table(xr_ctx* c, int k, int dim, int32_t* tab);
int xr_density_impl(xr_ctx* c, const char* ops, int n_orbs, int n_core, int n_elec_ket,
                    int64_t nI, int64_t nP, const double* zbra, int64_t nJ, int64_t nQ,
                    const double* zket, double* out);
