// Many independent meshes of uneven sizes advanced together.
//
// Every instance is a fused step (its own meshes, maps, impedance kernels,
// counters and histories), with the direct wake or the FFT wake.  The batch
// launches the step kernel once per step for all instances that share a
// kernel variant and workgroup size: the workgroups of one launch are the
// row tiles of all its instances (largest instances first), each reading
// its instance's descriptor, so instances of 128^2 to 1024^2 cells share
// the GPU in one grid.  Launches of different variants run as parallel
// branches.  The FFT-wake instances' field kernels (wake potential and CSR
// power, one workgroup per instance, each its own transform length) run as
// one launch before them.  The steps are captured into a hipGraph and
// replayed.
#include <algorithm>
#include <map>
#include <memory>
#include <numeric>
#include <tuple>

#include "common.hpp"
#include "step.hpp"
#include "step_kernel.hpp"

using namespace vfps_impl;

namespace {

struct Group {
    StepVariant variant{};
    uint32_t threads = 256;
    uint32_t lds_class = 0;         // LDS bytes of a tile, in 8 KiB units
    uint32_t maxn = 0;              // largest mesh of the group
    bool persist = false;           // grid sized to the resident workgroups
    uint32_t strip = 0;             // strip form (each member's
                                    // descriptor holds its tiles per strip)
    uint32_t grid = 0;              // workgroups (0: one per map entry)
    std::vector<uint32_t> members;  // batch indices
    InstDesc* descs_d = nullptr;
    uint32_t* blockmap_d = nullptr;
    uint32_t nblocks = 0;
    size_t lds = 0;
    hipStream_t stream = nullptr;   // branch stream (groups after the first)
    hipEvent_t done = nullptr;
};

// The field kernels of the FFT-wake members with one (wake, CSR) pair: a
// workgroup per member, ahead of the step kernels of the same step.
struct FieldLaunch {
    bool wake = false, csr = false;
    uint32_t ninst = 0, Mmax = 0;
    FieldParams* params_d = nullptr;
};

void group_free(Group& g)
{
    hipFree(g.descs_d);
    hipFree(g.blockmap_d);
    if (g.stream) hipStreamDestroy(g.stream);
    if (g.done) hipEventDestroy(g.done);
    g = Group{};
}

}  // namespace

struct vfps_batch {
    vfps_dev* dev = nullptr;
    std::vector<vfps_step*> members;
    std::vector<Group> groups;
    std::vector<FieldLaunch> fields;
    bool merge = false;   // one launch per variant, workgroup size and form
                          // (LDS classes merged: every tile gets the largest
                          // tile's LDS)
    hipEvent_t fork = nullptr;
    uint32_t parity = 0;
    // captured runs of k steps from parity p, key k << 1 | p
    std::map<uint32_t, std::pair<hipGraph_t, hipGraphExec_t>> graphs;
};

namespace {

int drop_graph(vfps_batch* b)
{
    for (auto& kv : b->graphs) {
        VFPS_HIP(hipGraphExecDestroy(kv.second.second));
        VFPS_HIP(hipGraphDestroy(kv.second.first));
    }
    b->graphs.clear();
    return VFPS_OK;
}

void free_groups(vfps_batch* b)
{
    for (auto& g : b->groups) group_free(g);
    b->groups.clear();
    for (auto& f : b->fields) hipFree(f.params_d);
    b->fields.clear();
}

// the FFT-wake members' field launches, one per (wake, CSR) pair
int build_fields(vfps_batch* b)
{
    for (int pair = 0; pair < 4; pair++) {
        const bool wake = (pair & 1) != 0, csr = (pair & 2) != 0;
        std::vector<FieldParams> ps;
        uint32_t Mmax = 0;
        for (const vfps_step* s : b->members) {
            if (s->direct || (s->desc.use_wake != 0) != wake ||
                (s->desc.use_csr != 0) != csr) {
                continue;
            }
            ps.push_back(s->fparams_h);
            Mmax = std::max(Mmax, s->M);
        }
        if (ps.empty()) continue;
        VFPS_CHECK_ARG(field_lds_bytes(Mmax) <= b->dev->lds_max,
                       "batch: an FFT length exceeds the LDS of one CU");
        FieldLaunch f;
        f.wake = wake;
        f.csr = csr;
        f.ninst = (uint32_t)ps.size();
        f.Mmax = Mmax;
        b->fields.push_back(f);
        FieldLaunch& fl = b->fields.back();
        VFPS_HIP(hipMalloc(&fl.params_d, sizeof(FieldParams) * ps.size()));
        VFPS_HIP(hipMemcpy(fl.params_d, ps.data(), sizeof(FieldParams) * ps.size(),
                           hipMemcpyHostToDevice));
    }
    return VFPS_OK;
}

bool same_variant(const StepVariant& a, const StepVariant& c)
{
    return a.scheme == c.scheme && a.it == c.it && a.ipfp == c.ipfp &&
           a.direct == c.direct && a.clamp == c.clamp;
}

// one group's device descriptors and block map
int build_group(vfps_batch* b, Group& g)
{
    // largest instances first, so their tiles are dealt out first
    std::vector<uint32_t> order(g.members.size());
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
        const auto* a = b->members[g.members[x]];
        const auto* c = b->members[g.members[y]];
        return (uint64_t)a->desc.n * a->desc.n > (uint64_t)c->desc.n * c->desc.n;
    });
    std::vector<uint32_t> map;
    g.lds = 0;
    g.maxn = 0;
    for (uint32_t k : order) {
        const vfps_step* s = b->members[g.members[k]];
        g.maxn = std::max(g.maxn, s->desc.n);
        const uint32_t units = g.strip ? s->strips() : s->tiles();
        for (uint32_t t = 0; t < units; t++) {
            map.push_back((k << 16) | t);
        }
        g.lds = std::max(g.lds, step_lds_of(s));
    }
    while (map.size() % 8u != 0u) {
        map.push_back(STEP_NO_TILE);
    }
    VFPS_CHECK_ARG(g.lds <= b->dev->lds_max,
                   "batch: a tile exceeds the LDS of one CU");
    VFPS_CHECK_ARG(!(g.persist && !g.variant.direct),
                   "batch: the resident-grid form needs the direct wake");
    g.nblocks = (uint32_t)map.size();
    g.grid = 0;
    VFPS_HIP(hipMalloc(&g.blockmap_d, sizeof(uint32_t) * map.size()));
    VFPS_HIP(hipMemcpy(g.blockmap_d, map.data(), sizeof(uint32_t) * map.size(),
                       hipMemcpyHostToDevice));
    // the members' descriptors (host copies), one upload
    std::vector<InstDesc> descs(g.members.size());
    for (size_t k = 0; k < g.members.size(); k++) {
        descs[k] = b->members[g.members[k]]->inst_h;
    }
    VFPS_HIP(hipMalloc(&g.descs_d, sizeof(InstDesc) * descs.size()));
    VFPS_HIP(hipMemcpy(g.descs_d, descs.data(), sizeof(InstDesc) * descs.size(),
                       hipMemcpyHostToDevice));
    int rc = g.strip ? strip_set_lds_limit(g.variant, g.lds)
                     : step_set_lds_limit(g.variant, g.lds);
    if (rc != VFPS_OK) return rc;
    if (g.persist && !g.strip) {
        // as many workgroups as stay resident on the device at once, a
        // multiple of 8 (one share per XCD), each walking several tiles
        int per_cu = 0;
        rc = step_occupancy(g.variant, g.threads, g.maxn, g.lds, &per_cu);
        if (rc != VFPS_OK) return rc;
        const uint32_t res = (uint32_t)std::max(per_cu, 1) *
                             (uint32_t)std::max(b->dev->ncu, 8);
        const uint32_t grid = std::min(g.nblocks, res) & ~7u;
        g.grid = grid < g.nblocks ? std::max(grid, 8u) : 0u;
    }
    VFPS_HIP(hipEventCreateWithFlags(&g.done, hipEventDisableTiming));
    return VFPS_OK;
}

// (re)build the launch groups from the members' current configuration:
// one launch per kernel variant, workgroup size and LDS size class (every
// workgroup of a launch gets the largest tile's LDS, so mixing small and
// large tiles would cost the small ones occupancy)
int build_groups(vfps_batch* b)
{
    int rc = drop_graph(b);
    if (rc != VFPS_OK) return rc;
    free_groups(b);
    for (uint32_t i = 0; i < b->members.size(); i++) {
        const vfps_step* s = b->members[i];
        const uint32_t cls = ceil_div(step_lds_of(s), 8192);
        Group* g = nullptr;
        for (auto& e : b->groups) {
            if (same_variant(e.variant, s->variant) && e.threads == s->threads &&
                (b->merge || e.lds_class == cls) && e.persist == s->persist &&
                (e.strip != 0) == (s->strip != 0)) {
                g = &e;
                break;
            }
        }
        if (!g) {
            b->groups.emplace_back();
            g = &b->groups.back();
            g->variant = s->variant;
            g->threads = s->threads;
            g->lds_class = cls;
            g->persist = s->persist;
            g->strip = s->strip;
        }
        g->members.push_back(i);
    }
    for (size_t k = 0; k < b->groups.size(); k++) {
        Group& g = b->groups[k];
        if ((rc = build_group(b, g)) != VFPS_OK) return rc;
        if (k > 0) {
            VFPS_HIP(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
        }
    }
    return build_fields(b);
}

// one step of every instance: the first group on the device stream, the
// others on branch streams forked from and joined back into it
int launch_step(vfps_batch* b, uint32_t parity)
{
    hipStream_t main = b->dev->stream;
    // the FFT-wake members' wake potentials, CSR powers and step counts
    for (const auto& f : b->fields) {
        int rc = launch_fields(f.wake, f.csr, f.params_d, f.ninst, f.Mmax,
                               parity, main);
        if (rc != VFPS_OK) return rc;
    }
    if (b->groups.size() > 1) {
        VFPS_HIP(hipEventRecord(b->fork, main));
    }
    for (size_t k = 0; k < b->groups.size(); k++) {
        Group& g = b->groups[k];
        hipStream_t st = k == 0 ? main : g.stream;
        if (k > 0) {
            VFPS_HIP(hipStreamWaitEvent(st, b->fork, 0));
        }
        int rc = g.strip
                     ? launch_strip_kernel(g.variant, g.descs_d, g.blockmap_d,
                                           g.nblocks, g.lds, g.threads, g.maxn,
                                           parity, st)
                     : launch_step_kernel(g.variant, g.descs_d, g.blockmap_d,
                                          g.nblocks, g.lds, g.threads, g.maxn,
                                          parity, st, nullptr, g.grid);
        if (rc != VFPS_OK) return rc;
        if (k > 0) {
            VFPS_HIP(hipEventRecord(g.done, st));
        }
    }
    for (size_t k = 1; k < b->groups.size(); k++) {
        VFPS_HIP(hipStreamWaitEvent(main, b->groups[k].done, 0));
    }
    return VFPS_OK;
}

// longest captured run; a run of nsteps replays it nsteps / BATCH_GRAPH
// times, then one captured run of the remaining steps
constexpr uint32_t BATCH_GRAPH = 64;

// the captured run of `steps` steps from parity p (captured on first use)
int batch_graph(vfps_batch* b, uint32_t steps, uint32_t p, hipGraphExec_t* out)
{
    const uint32_t key = steps << 1 | (p & 1u);
    auto it = b->graphs.find(key);
    if (it != b->graphs.end()) {
        *out = it->second.second;
        return VFPS_OK;
    }
    if (b->graphs.size() >= 16) {
        int rc = drop_graph(b);
        if (rc != VFPS_OK) return rc;
    }
    hipStream_t st = b->dev->stream;
    VFPS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    for (uint32_t k = 0; k < steps; k++) {
        int rc = launch_step(b, (p + k) & 1u);
        if (rc != VFPS_OK) {
            hipGraph_t g = nullptr;
            hipStreamEndCapture(st, &g);
            if (g) hipGraphDestroy(g);
            return rc;
        }
    }
    hipGraph_t def = nullptr;
    hipGraphExec_t exec = nullptr;
    VFPS_HIP(hipStreamEndCapture(st, &def));
    hipError_t e = hipGraphInstantiate(&exec, def, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        hipGraphDestroy(def);
        return fail(VFPS_ERR_HIP, hipGetErrorString(e));
    }
    b->graphs[key] = {def, exec};
    // uploaded now, so that the first replay does not pay for it
    VFPS_HIP(hipGraphUpload(exec, st));
    *out = exec;
    return VFPS_OK;
}

void sync_parity(vfps_batch* b)
{
    for (auto* s : b->members) s->parity = b->parity;
}

// Retile the members `pick` selects (persistent: a resident grid walking
// the tiles; strip: tiles per strip, 0 for a workgroup per tile).  On
// failure every member keeps its previous tiling.
template <typename Pick>
int retile_members(vfps_batch* b, Pick pick, uint32_t T, uint32_t threads,
                   bool persistent, uint32_t strip)
{
    VFPS_HIP(hipStreamSynchronize(b->dev->stream));
    sync_parity(b);
    struct Old {
        vfps_step* s;
        uint32_t T, thr, strip;
        bool p;
    };
    std::vector<Old> old;
    int rc = VFPS_OK;
    for (uint32_t i = 0; i < b->members.size(); i++) {
        vfps_step* s = b->members[i];
        if (!pick(i, s)) continue;
        old.push_back({s, s->T, s->threads, s->strip, s->persist});
        s->persist = persistent && strip == 0;
        s->strip = strip;
        if ((rc = step_retile(s, T, threads)) != VFPS_OK) break;
    }
    if (rc == VFPS_OK) {
        rc = build_groups(b);
    }
    if (rc != VFPS_OK) {
        const std::string msg = vfps_last_error();
        for (auto& o : old) {
            o.s->persist = o.p;
            o.s->strip = o.strip;
            step_retile(o.s, o.T, o.thr);
        }
        build_groups(b);
        return fail(rc, msg.c_str());
    }
    return VFPS_OK;
}

void batch_free(vfps_batch* b)
{
    if (!b) return;
    if (b->dev) hipStreamSynchronize(b->dev->stream);
    drop_graph(b);
    free_groups(b);
    if (b->fork) hipEventDestroy(b->fork);
    for (auto* s : b->members) vfps_step_destroy(s);
    delete b;
}

// state of one instance saved around the autotuner's timing runs
struct Snapshot {
    float* mesh = nullptr;
    float* proj = nullptr;
    uint32_t counter = 0;
};

}  // namespace

extern "C" {

int vfps_batch_create(vfps_dev* dev, uint32_t ninst,
                      const vfps_step_desc* descs, const float* const* inits,
                      vfps_impedance* const* zwake,
                      vfps_impedance* const* zcsr, vfps_batch** out)
{
    VFPS_CHECK_ARG(dev && descs && inits && out && ninst > 0,
                   "vfps_batch_create: NULL argument or empty batch");
    VFPS_CHECK_ARG(ninst <= 0xFFFFu, "vfps_batch_create: too many instances");
    VFPS_TRY_BEGIN
    auto* b = new vfps_batch();
    std::unique_ptr<vfps_batch, void (*)(vfps_batch*)> guard(b, batch_free);
    b->dev = dev;
    for (uint32_t i = 0; i < ninst; i++) {
        vfps_step_desc d = descs[i];
        d.fused = 1;
        d.wake_method = d.wake_method == 1 ? 1u : 2u;
        vfps_step* s = nullptr;
        int rc = vfps_step_create(dev, &d, inits[i], zwake ? zwake[i] : nullptr,
                                  zcsr ? zcsr[i] : nullptr, &s);
        if (rc != VFPS_OK) return rc;
        s->in_batch = true;
        b->members.push_back(s);
        VFPS_CHECK_ARG(s->tiles() <= 0xFFFFu, "vfps_batch_create: too many tiles");
    }
    VFPS_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    int rc = build_groups(b);
    if (rc != VFPS_OK) return rc;
    VFPS_HIP(hipStreamSynchronize(dev->stream));
    *out = guard.release();
    return VFPS_OK;
    VFPS_TRY_END
}

int vfps_batch_destroy(vfps_batch* b)
{
    batch_free(b);
    return VFPS_OK;
}

int vfps_batch_run(vfps_batch* b, uint32_t nsteps)
{
    VFPS_CHECK_ARG(b, "vfps_batch_run: NULL batch");
    hipStream_t st = b->dev->stream;
    uint32_t left = nsteps;
    int rc;
    // host launches when the members ask for them (no_graph)
    const bool host = !b->members.empty() && b->members[0]->desc.no_graph;
    while (host && left > 0) {
        if ((rc = launch_step(b, b->parity)) != VFPS_OK) return rc;
        b->parity ^= 1u;
        left--;
    }
    // captured runs: full ones of BATCH_GRAPH steps (even, so the parity
    // stays), then one of the remaining steps
    while (left > 0) {
        const uint32_t k = left >= BATCH_GRAPH ? BATCH_GRAPH : left;
        hipGraphExec_t g = nullptr;
        if ((rc = batch_graph(b, k, b->parity, &g)) != VFPS_OK) return rc;
        VFPS_HIP(hipGraphLaunch(g, st));
        left -= k;
        b->parity ^= (k & 1u);
    }
    if (nsteps > 0) {
        for (auto& g : b->groups) {
            // (the FFT form's field kernel has recorded each step's power)
            if (!g.variant.direct) continue;
            rc = launch_step_finalize(g.descs_d, (uint32_t)g.members.size(),
                                      b->parity, st);
            if (rc != VFPS_OK) return rc;
        }
    }
    // every member's error flag comes back in the same stream: one wait
    for (auto* s : b->members) {
        VFPS_HIP(hipMemcpyAsync(s->err_h, s->err, sizeof(uint32_t),
                                hipMemcpyDeviceToHost, st));
    }
    VFPS_HIP(hipStreamSynchronize(st));
    sync_parity(b);
    for (auto* s : b->members) {
        const uint32_t err = *s->err_h;
        if (err) {
            return fail(VFPS_ERR_INVALID, "batch step: a tap fell outside the "
                                          "halo rows (internal error)");
        }
    }
    return VFPS_OK;
}

int vfps_batch_prepare(vfps_batch* b, uint32_t nsteps)
{
    VFPS_CHECK_ARG(b, "vfps_batch_prepare: NULL batch");
    VFPS_TRY_BEGIN
    hipGraphExec_t g = nullptr;
    int rc;
    if (nsteps >= BATCH_GRAPH &&
        (rc = batch_graph(b, BATCH_GRAPH, b->parity, &g)) != VFPS_OK) {
        return rc;
    }
    if (nsteps % BATCH_GRAPH &&
        (rc = batch_graph(b, nsteps % BATCH_GRAPH, b->parity, &g)) != VFPS_OK) {
        return rc;
    }
    VFPS_HIP(hipStreamSynchronize(b->dev->stream));
    return VFPS_OK;
    VFPS_TRY_END
}

int vfps_batch_record(vfps_batch* b, uint32_t capacity, int want_profiles)
{
    VFPS_CHECK_ARG(b, "vfps_batch_record: NULL batch");
    for (auto* s : b->members) {
        int rc = vfps_step_record(s, capacity, want_profiles);
        if (rc != VFPS_OK) return rc;
    }
    return build_groups(b);
}

int vfps_batch_get_history(vfps_batch* b, uint32_t inst, uint32_t* count,
                           float* csr_power, float* profiles)
{
    VFPS_CHECK_ARG(b && inst < b->members.size(), "vfps_batch_get_history: "
                                                  "bad instance");
    return vfps_step_get_history(b->members[inst], count, csr_power, profiles);
}

int vfps_batch_get_mesh(vfps_batch* b, uint32_t inst, float* mesh,
                        float* proj, float* wake)
{
    VFPS_CHECK_ARG(b && inst < b->members.size(), "vfps_batch_get_mesh: "
                                                  "bad instance");
    sync_parity(b);
    return vfps_step_get_mesh(b->members[inst], mesh, proj, wake);
}

int vfps_batch_get_projections(vfps_batch* b, float* out)
{
    VFPS_CHECK_ARG(b && out, "vfps_batch_get_projections: NULL argument");
    hipStream_t st = b->dev->stream;
    VFPS_HIP(hipStreamSynchronize(st));
    sync_parity(b);
    size_t o = 0;
    for (auto* s : b->members) {
        const vfps_mesh* cur = (s->parity & 1u) ? s->t2 : s->t1;
        VFPS_HIP(hipMemcpyAsync(out + o, cur->proj, sizeof(float) * s->desc.n,
                                hipMemcpyDeviceToHost, st));
        o += s->desc.n;
    }
    VFPS_HIP(hipStreamSynchronize(st));
    return VFPS_OK;
}

int vfps_batch_get_csr_histories(vfps_batch* b, uint32_t first, uint32_t count,
                                 float* out)
{
    VFPS_CHECK_ARG(b && out, "vfps_batch_get_csr_histories: NULL argument");
    hipStream_t st = b->dev->stream;
    VFPS_HIP(hipStreamSynchronize(st));
    size_t o = 0;
    for (auto* s : b->members) {
        VFPS_CHECK_ARG(s->csr_hist && first + count <= s->hist_cap,
                       "vfps_batch_get_csr_histories: outside the recorded "
                       "capacity");
        VFPS_HIP(hipMemcpyAsync(out + o, s->csr_hist + first,
                                sizeof(float) * count, hipMemcpyDeviceToHost,
                                st));
        o += count;
    }
    VFPS_HIP(hipStreamSynchronize(st));
    return VFPS_OK;
}

int vfps_batch_info(vfps_batch* b, uint32_t* ninst, uint32_t* ngroups,
                    uint64_t* cells)
{
    VFPS_CHECK_ARG(b, "vfps_batch_info: NULL batch");
    uint64_t c = 0;
    for (auto* s : b->members) c += (uint64_t)s->desc.n * s->desc.n;
    if (ninst) *ninst = (uint32_t)b->members.size();
    if (ngroups) *ngroups = (uint32_t)b->groups.size();
    if (cells) *cells = c;
    return VFPS_OK;
}

int vfps_batch_launches(vfps_batch* b, uint32_t* out, uint32_t cap,
                        uint32_t* ngroups)
{
    VFPS_CHECK_ARG(b && (out || cap == 0), "vfps_batch_launches: bad argument");
    for (size_t k = 0; k < b->groups.size() && k < cap; k++) {
        const Group& g = b->groups[k];
        uint32_t* r = out + 4 * k;
        r[0] = g.threads;
        r[1] = g.nblocks;
        r[2] = g.grid ? g.grid : g.nblocks;
        r[3] = (uint32_t)g.lds;
    }
    if (ngroups) *ngroups = (uint32_t)b->groups.size();
    return VFPS_OK;
}

int vfps_batch_config(vfps_batch* b, uint32_t inst, uint32_t* out,
                      uint32_t nout)
{
    VFPS_CHECK_ARG(b && inst < b->members.size() && out && nout >= 6,
                   "vfps_batch_config: bad argument");
    return vfps_step_config(b->members[inst], out, nout);
}

int vfps_batch_retile(vfps_batch* b, uint32_t inst, uint32_t tile_rows,
                      uint32_t threads, int persistent)
{
    VFPS_CHECK_ARG(b && inst < b->members.size(), "vfps_batch_retile: bad "
                                                  "instance");
    VFPS_CHECK_ARG(tile_rows >= 1 && threads >= 64 && threads % 64 == 0 &&
                       threads <= STEP_MAX_THREADS,
                   "vfps_batch_retile: tile rows or workgroup size");
    VFPS_TRY_BEGIN
    return retile_members(
        b, [&](uint32_t i, const vfps_step*) { return i == inst; }, tile_rows,
        threads, persistent != 0, 0);
    VFPS_TRY_END
}

int vfps_batch_retile_strip(vfps_batch* b, uint32_t inst, uint32_t tile_rows,
                            uint32_t threads, uint32_t strip_tiles)
{
    VFPS_CHECK_ARG(b && inst < b->members.size(), "vfps_batch_retile_strip: "
                                                  "bad instance");
    VFPS_CHECK_ARG(tile_rows >= 1 && threads >= 64 && threads % 64 == 0 &&
                       threads <= STEP_MAX_THREADS && strip_tiles >= 1,
                   "vfps_batch_retile_strip: tile rows, workgroup size or "
                   "tiles per strip");
    VFPS_TRY_BEGIN
    return retile_members(
        b, [&](uint32_t i, const vfps_step*) { return i == inst; }, tile_rows,
        threads, false, strip_tiles);
    VFPS_TRY_END
}

int vfps_batch_set_merge(vfps_batch* b, int merge)
{
    VFPS_CHECK_ARG(b, "vfps_batch_set_merge: NULL batch");
    VFPS_TRY_BEGIN
    VFPS_HIP(hipStreamSynchronize(b->dev->stream));
    sync_parity(b);
    b->merge = merge != 0;
    return build_groups(b);
    VFPS_TRY_END
}

int vfps_batch_retile_size(vfps_batch* b, uint32_t n, uint32_t tile_rows,
                           uint32_t threads, int persistent)
{
    VFPS_CHECK_ARG(b, "vfps_batch_retile_size: NULL batch");
    VFPS_CHECK_ARG(tile_rows >= 1 && threads >= 64 && threads % 64 == 0 &&
                       threads <= STEP_MAX_THREADS,
                   "vfps_batch_retile_size: tile rows or workgroup size");
    VFPS_TRY_BEGIN
    return retile_members(
        b, [&](uint32_t, const vfps_step* s) { return s->desc.n == n; },
        tile_rows, threads, persistent != 0, 0);
    VFPS_TRY_END
}

int vfps_batch_retile_size_strip(vfps_batch* b, uint32_t n, uint32_t tile_rows,
                                 uint32_t threads, uint32_t strip_tiles)
{
    VFPS_CHECK_ARG(b, "vfps_batch_retile_size_strip: NULL batch");
    VFPS_CHECK_ARG(tile_rows >= 1 && threads >= 64 && threads % 64 == 0 &&
                       threads <= STEP_MAX_THREADS && strip_tiles >= 1,
                   "vfps_batch_retile_size_strip: tile rows, workgroup size "
                   "or tiles per strip");
    VFPS_TRY_BEGIN
    return retile_members(
        b, [&](uint32_t, const vfps_step* s) { return s->desc.n == n; },
        tile_rows, threads, false, strip_tiles);
    VFPS_TRY_END
}

int vfps_batch_timing(vfps_batch* b, uint32_t nsteps, float* ms_step)
{
    VFPS_CHECK_ARG(b && nsteps > 0 && ms_step, "vfps_batch_timing: bad argument");
    hipStream_t st = b->dev->stream;
    hipEvent_t e0, e1;
    VFPS_HIP(hipEventCreate(&e0));
    VFPS_HIP(hipEventCreate(&e1));
    double tot = 0;
    int rc = VFPS_OK;
    for (uint32_t k = 0; k < nsteps && rc == VFPS_OK; k++) {
        hipEventRecord(e0, st);
        rc = launch_step(b, b->parity);
        hipEventRecord(e1, st);
        hipEventSynchronize(e1);
        float ms = 0;
        hipEventElapsedTime(&ms, e0, e1);
        tot += ms;
        b->parity ^= 1u;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    sync_parity(b);
    if (rc != VFPS_OK) return rc;
    *ms_step = (float)(tot / nsteps);
    return VFPS_OK;
}

// Choose the tile height and workgroup size per mesh size: for each size
// class, every candidate (T, threads) is timed on that class's instances
// alone (one launch per step, HIP events); the fastest is kept.  The state
// of every instance is saved before and restored after, so tuning does not
// advance the simulation.  table rows: n, T, threads, microseconds per
// step, chosen (1/0).
int vfps_batch_autotune(vfps_batch* b, uint32_t steps, float* table,
                        uint32_t cap, uint32_t* nrows)
{
    VFPS_CHECK_ARG(b && steps > 0, "vfps_batch_autotune: bad argument");
    VFPS_TRY_BEGIN
    hipStream_t st = b->dev->stream;
    VFPS_HIP(hipStreamSynchronize(st));
    sync_parity(b);
    const uint32_t par = b->parity;
    // save the state
    std::vector<Snapshot> snap(b->members.size());
    for (size_t i = 0; i < b->members.size(); i++) {
        vfps_step* s = b->members[i];
        const size_t cells = (size_t)s->desc.n * s->desc.n;
        VFPS_HIP(hipMalloc(&snap[i].mesh, sizeof(float) * cells));
        VFPS_HIP(hipMalloc(&snap[i].proj, sizeof(float) * s->desc.n));
        const vfps_mesh* cur = par ? s->t2 : s->t1;
        VFPS_HIP(hipMemcpy(snap[i].mesh, cur->data, sizeof(float) * cells,
                           hipMemcpyDeviceToDevice));
        VFPS_HIP(hipMemcpy(snap[i].proj, cur->proj, sizeof(float) * s->desc.n,
                           hipMemcpyDeviceToDevice));
        VFPS_HIP(hipMemcpy(&snap[i].counter, s->counter + (par & 1u),
                           sizeof(uint32_t), hipMemcpyDeviceToHost));
    }
    std::map<uint32_t, std::vector<uint32_t>> classes;
    for (uint32_t i = 0; i < b->members.size(); i++) {
        classes[b->members[i]->desc.n].push_back(i);
    }
    const uint32_t even = steps + (steps & 1u);
    uint32_t rows = 0;
    int rc = VFPS_OK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    VFPS_HIP(hipEventCreate(&e0));
    VFPS_HIP(hipEventCreate(&e1));
    // time one tiling of one size class on its instances alone: us per step
    // (dup: a persistent grid as large as the map, i.e. the plain launch)
    auto time_config = [&](const std::vector<uint32_t>& idx, uint32_t T,
                           uint32_t thr, bool pers, uint32_t strip,
                           uint32_t nsteps, float* us, bool* dup) -> int {
        *dup = false;
        for (uint32_t i : idx) {
            b->members[i]->persist = pers;
            b->members[i]->strip = strip;
            int r0 = step_retile(b->members[i], T, thr, false);
            if (r0 != VFPS_OK) return r0;
        }
        vfps_batch sub;
        sub.dev = b->dev;
        sub.fork = b->fork;
        for (uint32_t i : idx) sub.members.push_back(b->members[i]);
        int r1 = build_groups(&sub);
        if (r1 == VFPS_OK && pers) {
            *dup = true;
            for (const auto& g : sub.groups) *dup = *dup && g.grid == 0;
        }
        float ms = 0;
        if (r1 == VFPS_OK && !*dup) {
            // warm up, then time
            for (uint32_t k = 0; k < 2 && r1 == VFPS_OK; k++) {
                r1 = launch_step(&sub, (par + k) & 1u);
            }
            hipEventRecord(e0, st);
            for (uint32_t k = 0; k < nsteps && r1 == VFPS_OK; k++) {
                r1 = launch_step(&sub, (par + k) & 1u);
            }
            hipEventRecord(e1, st);
            hipEventSynchronize(e1);
            hipEventElapsedTime(&ms, e0, e1);
        }
        free_groups(&sub);
        sub.fork = nullptr;
        sub.members.clear();
        *us = 1e3f * ms / (float)nsteps;
        return r1;
    };
    struct Cand {
        uint32_t T, thr;
        bool pers;
        uint32_t strip;
        float us;
        uint32_t row;
        bool refined;
    };
    // per class: the chosen tiling and the fastest per-tile one
    struct Choice {
        uint32_t n;
        Cand best, best_tile;
        bool has_tile;
    };
    std::vector<Choice> choices;
    for (auto& cl : classes) {
        const uint32_t n = cl.first;
        struct Orig {
            uint32_t T, thr, strip;
            bool pers;
        };
        std::vector<Orig> orig;
        for (uint32_t i : cl.second) {
            const vfps_step* m = b->members[i];
            orig.push_back({m->T, m->threads, m->strip, m->persist});
        }
        // strips exist for the kick-drift maps with the direct wake; below
        // 1024 rows a size class has too few strips to fill the GPU beside
        // the other classes of a batch (at 512 rows strips timed alone can
        // beat tiles, yet the whole ragged batch then runs slower), so the
        // tuner does not try them there
        bool direct = true;  // the resident-grid form needs every member
        for (uint32_t i : cl.second) direct = direct && b->members[i]->direct;
        const bool strips_ok =
            n >= 1024 &&
            strip_kernel_for(b->members[cl.second[0]]->variant, 1) != nullptr &&
            !b->members[cl.second[0]]->desc.fast_projection;
        std::vector<Cand> cands;
        for (uint32_t thr : {128u, 256u, 512u}) {
            if (ceil_div(n, thr) > (uint32_t)STEP_MAX_COLS) continue;
            // (T, persistent, tiles per strip): a workgroup per tile, a
            // resident grid, strips of 16, 32 or 64 rows
            std::vector<std::tuple<uint32_t, bool, uint32_t>> forms;
            for (uint32_t T : {1u, 2u, 4u, 8u, 16u}) {
                forms.emplace_back(T, false, 0u);
                if (direct) forms.emplace_back(T, true, 0u);
            }
            if (strips_ok) {
                for (uint32_t T : {2u, 4u, 8u}) {
                    for (uint32_t S : {16u, 32u, 64u}) {
                        if (S > T && S <= n) forms.emplace_back(T, false, S / T);
                    }
                }
            }
            for (const auto& f : forms) {
                const uint32_t T = std::get<0>(f);
                const bool pers = std::get<1>(f);
                const uint32_t strip = std::get<2>(f);
                if (T > n) continue;
                float us = 0;
                bool dup = false;
                const int r = time_config(cl.second, T, thr, pers, strip, even,
                                          &us, &dup);
                if (r == VFPS_ERR_INVALID) continue;  // does not fit
                if (r != VFPS_OK) {
                    rc = r;
                    break;
                }
                if (dup) continue;
                cands.push_back({T, thr, pers, strip, us, rows++, false});
            }
            if (rc != VFPS_OK) break;
        }
        if (rc != VFPS_OK) break;
        // the three fastest again over four times the steps: short timings
        // are noisy, and the choice should not follow the noise
        std::sort(cands.begin(), cands.end(),
                  [](const Cand& a, const Cand& c) { return a.us < c.us; });
        for (size_t k = 0; k < cands.size() && k < 3 && rc == VFPS_OK; k++) {
            bool dup = false;
            rc = time_config(cl.second, cands[k].T, cands[k].thr, cands[k].pers,
                             cands[k].strip, 4 * even, &cands[k].us, &dup);
            cands[k].refined = true;
        }
        if (rc != VFPS_OK) break;
        const Cand* best = nullptr;
        for (const auto& c : cands) {
            if (c.refined && (!best || c.us < best->us)) best = &c;
        }
        if (best) {
            Choice ch{n, *best, *best, false};
            for (const auto& c : cands) {
                if (!c.pers && (!ch.has_tile || c.us < ch.best_tile.us)) {
                    ch.best_tile = c;
                    ch.has_tile = true;
                }
            }
            choices.push_back(ch);
        }
        for (size_t j = 0; j < cl.second.size(); j++) {
            vfps_step* m = b->members[cl.second[j]];
            if (best) {
                m->persist = best->pers;
                m->strip = best->strip;
                rc = step_retile(m, best->T, best->thr, false);
            } else {
                m->persist = orig[j].pers;
                m->strip = orig[j].strip;
                rc = step_retile(m, orig[j].T, orig[j].thr, false);
            }
            if (rc != VFPS_OK) break;
        }
        if (table) {
            for (const auto& c : cands) {
                if (c.row >= cap) continue;
                float* r = table + VFPS_TUNE_COLS * (size_t)c.row;
                r[0] = (float)n;
                r[1] = (float)c.T;
                r[2] = (float)c.thr;
                r[3] = c.us;
                r[4] = (&c == best) ? 1.0f : 0.0f;
                r[5] = c.pers ? 1.0f : 0.0f;
                r[6] = c.refined ? 1.0f : 0.0f;
                r[7] = (float)c.strip;
            }
        }
        if (rc != VFPS_OK) break;
    }
    // A persistent launch holds every resident slot, so it can starve the
    // other size classes' launches, which run beside it: with several
    // classes, a persistent choice must also win on the whole batch.
    if (rc == VFPS_OK && choices.size() > 1) {
        auto time_all = [&](float* us) -> int {
            vfps_batch sub;
            sub.dev = b->dev;
            sub.fork = b->fork;
            sub.members = b->members;
            int r1 = build_groups(&sub);
            float ms = 0;
            if (r1 == VFPS_OK) {
                for (uint32_t k = 0; k < 2 && r1 == VFPS_OK; k++) {
                    r1 = launch_step(&sub, (par + k) & 1u);
                }
                hipEventRecord(e0, st);
                for (uint32_t k = 0; k < 4 * even && r1 == VFPS_OK; k++) {
                    r1 = launch_step(&sub, (par + k) & 1u);
                }
                hipEventRecord(e1, st);
                hipEventSynchronize(e1);
                hipEventElapsedTime(&ms, e0, e1);
            }
            free_groups(&sub);
            sub.fork = nullptr;
            sub.members.clear();
            *us = 1e3f * ms / (float)(4 * even);
            return r1;
        };
        for (auto& ch : choices) {
            if (!ch.best.pers || !ch.has_tile || rc != VFPS_OK) continue;
            float t_pers = 0, t_tile = 0;
            if ((rc = time_all(&t_pers)) != VFPS_OK) break;
            for (uint32_t i : classes[ch.n]) {
                b->members[i]->persist = false;
                b->members[i]->strip = ch.best_tile.strip;
                if ((rc = step_retile(b->members[i], ch.best_tile.T,
                                      ch.best_tile.thr, false)) != VFPS_OK)
                    break;
            }
            if (rc != VFPS_OK) break;
            if ((rc = time_all(&t_tile)) != VFPS_OK) break;
            if (t_pers < t_tile) {
                // back to the persistent choice
                for (uint32_t i : classes[ch.n]) {
                    b->members[i]->persist = true;
                    b->members[i]->strip = 0;
                    if ((rc = step_retile(b->members[i], ch.best.T,
                                          ch.best.thr, false)) != VFPS_OK)
                        break;
                }
            } else if (table) {
                // the chosen mark moves to the per-tile tiling
                if (ch.best.row < cap) {
                    table[VFPS_TUNE_COLS * (size_t)ch.best.row + 4] = 0.0f;
                }
                if (ch.best_tile.row < cap) {
                    table[VFPS_TUNE_COLS * (size_t)ch.best_tile.row + 4] = 1.0f;
                }
            }
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    // restore the state (every copy is attempted; the first failure is
    // reported): also what the retiles above left pending
    hipError_t re = hipStreamSynchronize(st);
    auto keep = [&re](hipError_t e) {
        if (re == hipSuccess) re = e;
    };
    for (size_t i = 0; i < b->members.size(); i++) {
        vfps_step* s = b->members[i];
        const size_t cells = (size_t)s->desc.n * s->desc.n;
        const vfps_mesh* cur = par ? s->t2 : s->t1;
        keep(hipMemcpy(cur->data, snap[i].mesh, sizeof(float) * cells,
                       hipMemcpyDeviceToDevice));
        keep(hipMemcpy(cur->proj, snap[i].proj, sizeof(float) * s->desc.n,
                       hipMemcpyDeviceToDevice));
        const uint32_t cc[2] = {snap[i].counter, snap[i].counter};
        keep(hipMemcpy(s->counter, cc, sizeof(cc), hipMemcpyHostToDevice));
        if (s->csr_valid) {
            // (the direct form's record of which step the CSR shares are
            // from; the FFT form has none)
            keep(hipMemset(s->csr_valid, 0xFF, sizeof(uint32_t) * 2));
        }
        keep(hipFree(snap[i].mesh));
        keep(hipFree(snap[i].proj));
        s->parity = par;
    }
    if (rc != VFPS_OK) return rc;
    if (re != hipSuccess) return fail(VFPS_ERR_HIP, hipGetErrorString(re));
    if (nrows) *nrows = rows;
    return build_groups(b);
    VFPS_TRY_END
}

}  // extern "C"
