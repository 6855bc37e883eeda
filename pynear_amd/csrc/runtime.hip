// Library-wide runtime pieces of the C-ABI: the thread-local error string, the library stream,
// version and device queries.
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"

namespace pn {

static thread_local std::string g_last_error;

void set_last_error(const std::string &msg) { g_last_error = msg; }

hipStream_t lib_stream() {
    // One non-blocking stream per device, created on first use at default priority.
    static std::mutex mu;
    static hipStream_t streams[64] = {};
    int dev = current_device();
    PN_REQUIRE(dev >= 0 && dev < 64, PN_EDEVICE, "device id out of range");
    std::lock_guard<std::mutex> lk(mu);
    if (!streams[dev]) PN_HIP(hipStreamCreateWithFlags(&streams[dev], hipStreamNonBlocking));
    return streams[dev];
}

// Tuning table. Defaults are the measured best on MI355X for the workloads in bench.py.
static std::mutex g_tune_mu;
static std::map<std::string, int64_t> &tune_table() {
    static std::map<std::string, int64_t> t;
    static bool init = false;
    if (!init) {
        init = true;
        t["hamming.wg_per_cu"] = 8;   // scan workgroups per CU (4 waves each)
        t["hamming.qpl"] = 1;         // queries per lane in the Hamming scan (1 or 2)
        t["float.wg_per_cu"] = 16;
        t["merge.wave_per_query"] = 1; // 1: one wave merges one query's slice lists
        t["l2.mfma"] = 1;              // exact L2 (k <= 32) through the MFMA candidate scan + rerank
        t["l2.wg_per_cu"] = 2;
        t["l2.mfma_terms"] = 3;        // 3: hi.hi + hi.lo + lo.hi products; 1: hi.hi only
        t["l2.matrix"] = 1;            // exact L2 over <= 4096 rows of d > 64: key matrix + selection
        t["ivfl2.wg_per_cu"] = 2;      // IVF-L2 list scan: persistent workgroups per CU
        t["l2.wide"] = 1;             // exact L2 scan for dp = 128, 192, 256: queries held in registers
        t["l2.fuse_rerank"] = 1;      // exact L2: merge and rerank in one launch (candidates in registers)
        t["merge.lean"] = 1;          // wave merges without private lists: 1 for capacities >= 16, 3 always, 2 rounds only
        t["l2.rerank_reg"] = 2;       // exact L2 rerank: query chains in registers (1 on, 0 off, 2 for <= 1024 queries)
        t["l2.matrix_terms"] = 3;     // key-matrix path (<= 4096 wide rows): bf16 products per term (1 or 3)
        t["l2.kl4"] = 1;              // exact L2 scan: per-lane lists of 4 for k <= 6 (with >= 4 slices)
        t["l2.slice_major"] = 1;      // exact L2 scan grid: slices along x (rows of a slice stay in one XCD's L2)
        t["ivfl2.xcd_ranges"] = 0;     // 1: one item range per XCD; 0: one queue for all workgroups
        t["ivfl2.fp8_wg_per_cu"] = 1;  // fp8 scan: workgroups per CU
        t["ivfl2.bound"] = 1;          // fp8: per-query thresholds from the first probe's segment bounds
        t["ivfl2.prune"] = 1;          // fp8 with bounds: skip row groups that no query can need
        t["ivfl2.residual_xcd_np"] = 16; // fp8 residuals: per-XCD centroid split from this many probes per query
        t["ivfl2.two_pass"] = 0;       // fp8 without bounds: first probe, then the rest under its K2-th candidate
        t["ivfl2.fp8_iq2"] = 128;      // fp8 second pass of two: queries per work item (32 .. 128)
                                       // (the scan gains less than the second bucketing costs)
        t["ivfl2.fp8_ahead"] = 3;      // fp8 scan: fragments loaded 2 or 3 steps ahead
        t["ivfl2.k2"] = 40;            // candidates reranked in fp32 per query (max(4 k, this), <= 64)
        t["ivfl2.pass_margin_pct"] = 0; // > 0: scan the first probe, then the others bounded by its
                                         // k-th approximation x (1 + pct / 100); 0: one pass
        t["ivfbin.list_major"] = 1;    // IVF-binary fine search: 1 list-major scan + merge, 0 query-major
        t["batch.slice_rows"] = 2048;  // batched exact search: target rows per slice (<= 4096)
        t["batch.wg_per_cu"] = 2;      // batched exact search: persistent scan workgroups per CU
        t["batch.order"] = 2;          // batched exact search: 1 = tiles by decreasing cost, 2 = by slice, then cost
        t["batch.concurrent_scans"] = 1; // batched exact search: the three scan classes as parallel graph branches
        t["batch.share"] = 1;          // batched exact search: the slices of a query share their list bound
        t["batch.pipe"] = 3;           // batched exact search: software-pipelined scans (1: wide rows, 2: narrow rows)
        if (const char *env = std::getenv("PN_TUNE")) {
            std::string e(env);
            size_t pos = 0;
            while (pos < e.size()) {
                size_t comma = e.find(',', pos);
                std::string item = e.substr(pos, comma == std::string::npos ? std::string::npos : comma - pos);
                size_t eq = item.find('=');
                if (eq != std::string::npos) t[item.substr(0, eq)] = std::strtoll(item.c_str() + eq + 1, nullptr, 10);
                if (comma == std::string::npos) break;
                pos = comma + 1;
            }
        }
    }
    return t;
}

int64_t tune_get(const char *name) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto &t = tune_table();
    auto it = t.find(name);
    return it == t.end() ? 0 : it->second;
}

// Stage timer: begin/end events per launch, resolved by pn_stage_times().
static std::mutex g_stage_mu;
static bool g_stage_on = false;
struct StageRec {
    int stage;
    hipEvent_t b, e;
};
static std::vector<StageRec> g_stage_recs;
static hipEvent_t g_open_ev = nullptr;
static int g_open_stage = -1;
static int g_depth = 0; // nested scopes belong to the outermost open stage

void stage_mark(hipStream_t st, int stage, bool begin) {
    if (!g_stage_on) return;
    std::lock_guard<std::mutex> lk(g_stage_mu);
    if (begin) {
        if (g_depth++ > 0) return;
        hipEvent_t ev;
        if (hipEventCreate(&ev) != hipSuccess) return;
        (void)hipEventRecord(ev, st);
        g_open_ev = ev;
        g_open_stage = stage;
        return;
    }
    if (g_depth == 0 || --g_depth > 0) return;
    if (!g_open_ev) return;
    hipEvent_t ev;
    if (hipEventCreate(&ev) != hipSuccess) return;
    (void)hipEventRecord(ev, st);
    g_stage_recs.push_back({g_open_stage, g_open_ev, ev});
    g_open_ev = nullptr;
    g_open_stage = -1;
}

} // namespace pn

extern "C" pn_status pn_tune_set(const char *name, int64_t value) {
    return pn::guard([&] {
        PN_REQUIRE(name, PN_EINVAL, "name is null");
        std::lock_guard<std::mutex> lk(pn::g_tune_mu);
        auto &t = pn::tune_table();
        PN_REQUIRE(t.count(name), PN_EINVAL, std::string("unknown tuning parameter: ") + name);
        t[name] = value;
    });
}

extern "C" pn_status pn_tune_get(const char *name, int64_t *value) {
    return pn::guard([&] {
        PN_REQUIRE(name && value, PN_EINVAL, "null argument");
        std::lock_guard<std::mutex> lk(pn::g_tune_mu);
        auto &t = pn::tune_table();
        PN_REQUIRE(t.count(name), PN_EINVAL, std::string("unknown tuning parameter: ") + name);
        *value = t[name];
    });
}

extern "C" pn_status pn_stage_timing(int32_t enable) {
    return pn::guard([&] {
        std::lock_guard<std::mutex> lk(pn::g_stage_mu);
        pn::g_stage_on = enable != 0;
    });
}

extern "C" pn_status pn_stage_times(double *ms, int64_t *launches) {
    return pn::guard([&] {
        PN_REQUIRE(ms && launches, PN_EINVAL, "null argument");
        std::lock_guard<std::mutex> lk(pn::g_stage_mu);
        for (int s = 0; s < pn::STAGE_COUNT; ++s) {
            ms[s] = 0.0;
            launches[s] = 0;
        }
        for (auto &r : pn::g_stage_recs) {
            PN_HIP(hipEventSynchronize(r.e));
            float t = 0.f;
            PN_HIP(hipEventElapsedTime(&t, r.b, r.e));
            ms[r.stage] += t;
            launches[r.stage] += 1;
            (void)hipEventDestroy(r.b);
            (void)hipEventDestroy(r.e);
        }
        pn::g_stage_recs.clear();
    });
}

extern "C" const char *pn_last_error(void) { return pn::g_last_error.c_str(); }

extern "C" const char *pn_version(void) { return "pynear_amd 0.1.0 (gfx950)"; }

extern "C" pn_status pn_device_count(int32_t *count) {
    return pn::guard([&] {
        PN_REQUIRE(count, PN_EINVAL, "count is null");
        int c = 0;
        hipError_t e = hipGetDeviceCount(&c);
        *count = e == hipSuccess ? c : 0;
    });
}
