// Fused DUGKS step kernels (see dugks_fused.h).
#include <hip/hip_runtime.h>

#include "dugks_fused.h"

namespace dugks {

bool fused_plan_build(FusedPlan &p, const Geo &, const dugks_mesh *, const HostPrep &, const dugks_vset *, size_t,
                      std::string &err)
{
    p.ready = false;
    err = "fused path not built in this version";
    return false;
}
bool fused_plan_upload(FusedPlan &, Geo &, hipStream_t, std::string &err)
{
    err = "fused path not built in this version";
    return false;
}
bool fused_init_boundary(FusedPlan &, const Geo &, hipStream_t, std::string &err)
{
    err = "fused path not built in this version";
    return false;
}
void fused_plan_free(FusedPlan &p)
{
    for (void *q : p.allocs)
        if (q) (void)hipFree(q);
    p.allocs.clear();
}

template <class S>
void Fused<S>::face_pass(const Geo &, FusedPlan &, const S *, const S *, const Macros &, const double *, hipStream_t) {}
template <class S>
void Fused<S>::face_macros(const Geo &, FusedPlan &, Macros, double *, const double *, double *, hipStream_t) {}
template <class S>
void Fused<S>::update_pass(const Geo &, FusedPlan &, S *, S *, Macros, const Macros &, const double *, const double *,
                           hipStream_t) {}
template <class S>
void Fused<S>::cell_macros(const Geo &, FusedPlan &, Macros, const double *, hipStream_t) {}

template struct Fused<float>;
template struct Fused<double>;

}  // namespace dugks
