// Batched multi-source runs (one bit per source packed per vertex).
#include "engine.hpp"

struct mg_batch { int kind; };

extern "C" int mg_batch_build(const mg_graph *, int, int, const mg_options *, mg_batch **out) {
    if (out) *out = nullptr;
    return mg_fail(MG_ERR_STATE, "batched runs are not built yet");
}
extern "C" int mg_batch_run(mg_batch *, int, const int32_t *, float *) { return mg_fail(MG_ERR_STATE, "batched runs are not built yet"); }
extern "C" int mg_batch_get(mg_batch *, int, void *, int64_t) { return mg_fail(MG_ERR_STATE, "batched runs are not built yet"); }
extern "C" int mg_batch_release(mg_batch *b) { delete b; return MG_OK; }
