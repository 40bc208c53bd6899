// Ragged batches of independent af_convolve2 calls (project extension; the
// reference has no counterpart: each problem follows
// src/api/c/convolve.cpp:228-241 with AF_CONV_SPATIAL and AF_BATCH_NONE).
//
// A plan validates the problems, allocates their outputs, builds a device
// table of problems and a cost-sorted list of (problem, tile) work items,
// lets the autotuner pick the tile variant, and captures the single
// persistent launch in a hipGraph; running the plan replays the graph.
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/af/array.h"
#include "../../include/af/hip.h"
#include "../../include/afhip_kernels.h"
#include "autotune.hpp"
#include "runtime.hpp"

namespace afhip {

// Device layout of one problem; must match RaggedProblem in
// kernels/ragged.hip.
struct HostProblem {
    const float *in;
    float *out;
    const float *filt;
    int S0, S1, O0, O1;
    int FX, FY, sx, sy;
    int nbx, nby;
};

struct RaggedPlan {
    std::vector<af_array> outs;               // owned handles (caller gets retains)
    std::vector<std::shared_ptr<Buffer>> keep;  // inputs' and filters' storage
    std::shared_ptr<Buffer> probs, items;
    std::vector<HostProblem> hp;
    long long nitems[3] = {0, 0, 0};
    std::shared_ptr<Buffer> items_v[3];
    int variant = 0, grid = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~RaggedPlan() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (af_array a : outs) af_release_array(a);
    }
};

static constexpr int kRaggedMaxF = 15;

// Builds the work list of one tile variant: problems in descending per-tile
// cost (taps x tile area), all tiles of a problem together.
static std::shared_ptr<Buffer> build_items(RaggedPlan &P, int variant, long long &count) {
    int bx = 0, by = 0;
    check_k(afhip_k_ragged_tile(variant, &bx, &by), "ragged tile");
    std::vector<HostProblem> hp = P.hp;
    for (auto &q : hp) {
        q.nbx = (q.O0 + bx - 1) / bx;
        q.nby = (q.O1 + by - 1) / by;
    }
    std::vector<int> order(hp.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return hp[a].FX * hp[a].FY > hp[b].FX * hp[b].FY;
    });
    std::vector<int> it;
    for (int p : order)
        for (int t = 0; t < hp[p].nbx * hp[p].nby; ++t) {
            it.push_back(p);
            it.push_back(t);
        }
    count = (long long)it.size() / 2;
    auto buf = scratch(std::max<size_t>(it.size() * sizeof(int), 16));
    if (!it.empty())
        check_hip(hipMemcpyAsync(buf->ptr, it.data(), it.size() * sizeof(int), hipMemcpyHostToDevice,
                                 stream()),
                  "ragged items");
    // per-variant problem table (tile counts differ)
    auto pb = scratch(std::max<size_t>(hp.size() * sizeof(HostProblem), 16));
    check_hip(hipMemcpyAsync(pb->ptr, hp.data(), hp.size() * sizeof(HostProblem), hipMemcpyHostToDevice,
                             stream()),
              "ragged problems");
    sync_stream();  // host vectors go out of scope
    P.keep.push_back(pb);
    return std::shared_ptr<Buffer>(buf);
}

}  // namespace afhip

using namespace afhip;

struct afhip_ragged_plan_t {
    RaggedPlan p;
    std::shared_ptr<Buffer> probs_v[3];
};

extern "C" {

AFAPI af_err afhip_ragged_plan_create(afhip_ragged_plan *plan, af_array *outputs,
                                      const af_array *signals, const af_array *filters,
                                      const af_conv_mode *modes, unsigned count) {
    try {
        ARG_ASSERT(0, plan != nullptr);
        ARG_ASSERT(1, outputs != nullptr || count == 0);
        ARG_ASSERT(2, signals != nullptr || count == 0);
        ARG_ASSERT(3, filters != nullptr || count == 0);
        ARG_ASSERT(4, modes != nullptr || count == 0);
        auto *h = new afhip_ragged_plan_t();
        RaggedPlan &P = h->p;
        try {
            for (unsigned i = 0; i < count; ++i) {
                const ArrayImpl &s = get(signals[i]);
                const ArrayImpl &f = get(filters[i]);
                if (s.type != f32) TYPE_ERROR(2, s.type);
                // a 2-D signal: af_convolve2 then stays 2-D (or, for a filter
                // of one column, takes the equivalent LHS 1-D route,
                // src/api/c/convolve.cpp:231-234)
                ARG_ASSERT(2, s.dims.ndims() == 2);
                ARG_ASSERT(3, f.dims[2] == 1 && f.dims[3] == 1 && f.dims.ndims() >= 1);
                ARG_ASSERT(3, f.dims[0] <= kRaggedMaxF && f.dims[1] <= kRaggedMaxF);
                ARG_ASSERT(4, modes[i] == AF_CONV_DEFAULT || modes[i] == AF_CONV_EXPAND);
                std::shared_ptr<Buffer> fb = f.buf;
                if (f.type != f32) {
                    fb = scratch((size_t)f.elements() * 4);
                    check_k(afhip_k_cast(f32, fb->ptr, f.type, f.data(), f.elements(), stream()), "cast");
                }
                const bool ex = modes[i] == AF_CONV_EXPAND;
                // output dims: src/backend/cpu/convolve.cpp:44-58 with AF_BATCH_NONE
                const dim4 od = ex ? dim4(s.dims[0] + f.dims[0] - 1, s.dims[1] + f.dims[1] - 1) : s.dims;
                af_array o = make_handle(f32, od);
                P.outs.push_back(o);
                P.keep.push_back(s.buf);
                P.keep.push_back(fb);
                HostProblem q;
                q.in = (const float *)s.data();
                q.out = (float *)get(o).data();
                q.filt = (const float *)fb->ptr;
                q.S0 = (int)s.dims[0];
                q.S1 = (int)s.dims[1];
                q.O0 = (int)od[0];
                q.O1 = (int)od[1];
                q.FX = (int)f.dims[0];
                q.FY = (int)f.dims[1];
                q.sx = ex ? 0 : q.FX / 2;
                q.sy = ex ? 0 : q.FY / 2;
                q.nbx = q.nby = 0;
                P.hp.push_back(q);
            }
            std::string key = "ragged_f32:" + std::to_string(count);
            unsigned long long hsh = 1469598103934665603ull;
            for (const auto &q : P.hp) {
                const int v[6] = {q.S0, q.S1, q.FX, q.FY, q.sx, q.sy};
                for (int x : v) hsh = (hsh ^ (unsigned long long)x) * 1099511628211ull;
            }
            key += ":" + std::to_string(hsh);
            for (int v = 0; v < 3; ++v) {
                h->p.items_v[v] = build_items(P, v, P.nitems[v]);
                h->probs_v[v] = P.keep.back();
            }
            // Persistent grid: 8 workgroups per CU. AFHIP_RAGGED_GRID caps it
            // lower (each workgroup then walks more items of the list).
            P.grid = 256 * 8;
            if (const char *ge = std::getenv("AFHIP_RAGGED_GRID")) {
                const long long gv = std::atoll(ge);
                if (gv > 0 && gv < P.grid) P.grid = (int)gv;
            }
            const int choice = autotune_pick(key, 3, [&](int v) {
                check_k(afhip_k_convolve2_ragged(v, h->probs_v[v]->ptr, P.items_v[v]->ptr, P.nitems[v],
                                                 (int)std::min<long long>(P.grid, std::max(1LL, P.nitems[v])),
                                                 stream()),
                        "ragged");
            });
            P.variant = choice;
            // capture the persistent launch (nothing to capture without work)
            hipStream_t st = stream();
            if (P.nitems[choice] == 0) {
                for (unsigned i = 0; i < count; ++i) af_retain_array(&outputs[i], P.outs[i]);
                *plan = h;
                return AF_SUCCESS;
            }
            sync_stream();
            check_hip(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
            const int rc = afhip_k_convolve2_ragged(
                choice, h->probs_v[choice]->ptr, P.items_v[choice]->ptr, P.nitems[choice],
                (int)std::min<long long>(P.grid, std::max(1LL, P.nitems[choice])), st);
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(st, &g);
            check_k(rc, "ragged capture");
            check_hip(ec, "hipStreamEndCapture");
            P.graph = g;
            check_hip(hipGraphInstantiate(&P.exec, g, nullptr, nullptr, 0), "hipGraphInstantiate");
        } catch (...) {
            delete h;
            throw;
        }
        for (unsigned i = 0; i < count; ++i) af_retain_array(&outputs[i], P.outs[i]);
        *plan = h;
    }
    CATCHALL;
    return AF_SUCCESS;
}

AFAPI af_err afhip_ragged_plan_run(afhip_ragged_plan plan) {
    try {
        ARG_ASSERT(0, plan != nullptr);
        if (plan->p.exec) check_hip(hipGraphLaunch(plan->p.exec, stream()), "hipGraphLaunch");
    }
    CATCHALL;
    return AF_SUCCESS;
}

AFAPI af_err afhip_ragged_plan_variant(int *variant, const afhip_ragged_plan plan) {
    try {
        ARG_ASSERT(0, variant != nullptr);
        ARG_ASSERT(1, plan != nullptr);
        *variant = plan->p.variant;
    }
    CATCHALL;
    return AF_SUCCESS;
}

AFAPI af_err afhip_ragged_plan_release(afhip_ragged_plan plan) {
    try {
        if (plan) {
            sync_stream();
            delete plan;
        }
    }
    CATCHALL;
    return AF_SUCCESS;
}

AFAPI af_err afhip_convolve2_ragged(af_array *outputs, const af_array *signals,
                                   const af_array *filters, const af_conv_mode *modes,
                                   unsigned count) {
    afhip_ragged_plan plan = nullptr;
    af_err e = afhip_ragged_plan_create(&plan, outputs, signals, filters, modes, count);
    if (e != AF_SUCCESS) return e;
    e = afhip_ragged_plan_run(plan);
    const af_err e2 = afhip_ragged_plan_release(plan);
    return e != AF_SUCCESS ? e : e2;
}

}  // extern "C"
