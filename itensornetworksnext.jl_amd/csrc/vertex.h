// vertex.h -- vertex-centric BP path (placeholder until the MFMA kernel lands).
#pragma once
#include <string>
#include "common.h"

struct VertexPlan {
    std::string why = "vertex path not built";
};

static int vertex_plan_build(VertexPlan *p, int, int, int, const int *, const int *, const int *,
                             const int *, const long long *, const long long *, const int *,
                             const int *) {
    p->why = "vertex path not built";
    return 4;
}
static void vertex_plan_release(VertexPlan *) {}
static int vertex_upload_sites(VertexPlan *, hipStream_t, const double *, void **) { return 0; }
static int vertex_sweep(VertexPlan *, hipStream_t, DevTopo, const void *, const void *, void *,
                        const int *, double *, double *) {
    return 4;
}
