// Fused DUGKS step: two gather passes over the distributions per step (face moments,
// then flux + update + cell moments) instead of materialised face distributions.
#pragma once

#include <string>

#include "../../include/dugks.h"
#include "dugks_internal.h"

namespace dugks {

struct FusedPlan {
    bool ready = false;
    std::vector<void *> allocs;
};

bool fused_plan_build(FusedPlan &p, const Geo &g, const dugks_mesh *mesh, const HostPrep &prep,
                      const dugks_vset *vset, size_t es, std::string &err);
bool fused_plan_upload(FusedPlan &p, Geo &g, hipStream_t st, std::string &err);
bool fused_init_boundary(FusedPlan &p, const Geo &g, hipStream_t st, std::string &err);
void fused_plan_free(FusedPlan &p);

template <class S>
struct Fused {
    static void face_pass(const Geo &g, FusedPlan &p, const S *gT, const S *hT, const Macros &cm, const double *dt,
                          hipStream_t st);
    static void face_macros(const Geo &g, FusedPlan &p, Macros fm, double *rho_b, const double *dt, double *cfl_part,
                            hipStream_t st);
    static void update_pass(const Geo &g, FusedPlan &p, S *gT, S *hT, Macros cm, const Macros &fm,
                            const double *rho_b, const double *dt, hipStream_t st);
    static void cell_macros(const Geo &g, FusedPlan &p, Macros cm, const double *dt, hipStream_t st);
};

}  // namespace dugks
