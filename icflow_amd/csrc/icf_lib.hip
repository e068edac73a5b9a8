// icf_lib.hip -- libicflow_hip.so: the C ABI declared in include/icflow_hip.h.
//
// Single translation unit: the kernels of icf_*.inc.h share one __constant__ DevConsts.
// Device memory for staging records and slot maps is owned by the context and grows on
// demand; the caller's buffers (nodes, pcsr arrays, BCSR, rhs) are never retained.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>
#include <string.h>
#include <math.h>
#include <climits>
#include <string>
#include <vector>
#include <list>
#include <type_traits>
#include <atomic>
#include <algorithm>
#include <mutex>

#include "../../include/icflow_hip.h"

#include "icf_device.h"
namespace icf {
__constant__ DevConsts g_c;
}
#include "icf_elem.inc.h"
#include "icf_cz.inc.h"
#include "icf_cz_phase1.inc.h"
#include "icf_contact.inc.h"
#include "icf_contact_resp.inc.h"
#include "icf_reduce.inc.h"
#include "icf_fused.inc.h"
#include "icf_batch.inc.h"
#include "icf_nodes.inc.h"
#include "icf_pattern.inc.h"
#include "icf_broad.inc.h"

using namespace icf;

namespace {

// Device bytes a call reads or writes: up to MAXS spans, or `all` (conflicts with everything).
struct Span {
    uintptr_t lo, hi;
    bool w;
};
struct Access {
    static constexpr int MAXS = 12;
    Span s[MAXS];
    int n = 0;
    bool all = false;
    Access &add(const void *p, size_t bytes, bool w)
    {
        if (!p || !bytes) return *this;
        if (n == MAXS) {
            all = true;
            return *this;
        }
        s[n++] = {(uintptr_t)p, (uintptr_t)p + bytes, w};
        return *this;
    }
    template <typename T> Access &rd(const T *p, size_t count) { return add(p, count * sizeof(T), false); }
    template <typename T> Access &wr(const T *p, size_t count) { return add(p, count * sizeof(T), true); }
    // the whole allocation holding p, for buffers whose used extent the call does not know
    Access &alloc(const void *p, bool w)
    {
        if (!p) return *this;
        void *base = nullptr;
        size_t sz = 0;
        if (hipMemGetAddressRange(&base, &sz, const_cast<void *>(p)) == hipSuccess && sz) return add(base, sz, w);
        (void)hipGetLastError();
        all = true;
        return *this;
    }
    static Access everything()
    {
        Access a;
        a.all = true;
        return a;
    }
};

bool conflict(const Access &a, const Access &b)
{
    if (a.all || b.all) return true;
    for (int i = 0; i < a.n; i++)
        for (int j = 0; j < b.n; j++)
            if ((a.s[i].w || b.s[j].w) && a.s[i].lo < b.s[j].hi && b.s[j].lo < a.s[i].hi) return true;
    return false;
}

// A call issued on one lane, with the event recorded after its last launch.
struct LaneRec {
    Access a;
    hipEvent_t ev;
};
constexpr int LANE_EVENTS = 32, LANE_LOG = 24;   // a logged event is one of the lane's last LANE_LOG

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

enum Target { TGT_VALS = 0, TGT_RHS = 1, TGT_FORCE = 2 };

struct SrcKey {
    uint32_t kind;
    const int *base;
    int n, stride, field0, nq;
    bool operator==(const SrcKey &o) const
    {
        return kind == o.kind && base == o.base && n == o.n && stride == o.stride && field0 == o.field0 && nq == o.nq;
    }
};

struct SlotMap {
    int target = 0;
    int fixed_nkeys = -1;          // -1: number of keys = largest key + 1
    std::vector<SrcKey> srcs;
    int nkeys = 0;
    int ncodes = 0;
    int *off = nullptr;            // nkeys + 1
    uint32_t *codes = nullptr;     // ncodes
    int *keys = nullptr;           // accumulate mode: the keys with contributions (null: all keys)
    int nlaunch = 0;               // keys a reduction visits
    // BCSR-value maps, per candidate chunk shape c (cand_keys(c), cand_stage(c)) and staging
    // type (0: FP32 values, 1: FP64 values): chunk starts (list position, code position)
    int2 *chunks[NCAND][2] = {};
    int nchunks[NCAND][2] = {};
    int id = 0;                    // unique per map built
    // rhs maps of the fused call: the force map (by id) whose per-node tet sums seed the rows,
    // the outcome of that check, and per row the node and the first non-tet code
    int tsum_with = 0;
    bool tsum_ok = false;
    int *row_node = nullptr, *cz_begin = nullptr;
    // ... and per node of the force map its rhs row (-1: none), when every row has one tet node
    // (then the node-force pass also writes the rhs rows)
    int *node_row = nullptr;
    bool merge_ok = false;
    // copy of the key fields the map was built from; after a write that may have touched them
    // (suspect) the map is kept if they still compare equal
    int *snap = nullptr;
    bool suspect = false;
};

std::atomic<int> g_next_ctx_id{1};
// All contexts on a device share the one __constant__ g_c.  g_const_owner[d] is the id of the
// context whose constants it holds; g_dev_mu[d] is held by every call that launches kernels
// reading g_c, from the check of the owner to its last launch, so that another context
// cannot replace the constants in between (ConstUse).
int g_const_owner[64];
std::mutex g_dev_mu[64];

// Scratch of the allocation-free entry point (icf_find_closest_face_noalloc): one array per
// GPU, reserved when the code object loads on that GPU and held until the process ends.
// Host code never passes `scratch` itself to a kernel or a copy (on the host the name is a
// stand-in, not a device address): it uses the address that hipGetSymbolAddress returned for
// the device, looked up once per device by icf_create and kept in g_scratch_ptr.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
std::mutex g_scratch_mu;
void *g_scratch_ptr[64];           // per device: device address of `scratch`, null until looked up

// With `device` current: the device address of its scratch array (looked up on first use).
hipError_t scratch_for_device(int device, void **out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

}  // namespace

// Launch-shape roles besides the value reductions (icf_set_shape / icf_get_shape): the
// workgroup size (64, 128 or 256 lanes) of phase 1 of the fused call, of the narrow phase and
// of the impacts' phase 1, the lanes per (node, tet) tuple of the closest face (4, 8 or 16),
// and the order of the fused value launch's two phases (0: node forces first, then the BCSR
// values; 1: interleaved in rounds of eight workgroups).  Every candidate gives the same
// results (each unit's and key's work does not depend on the shape).
constexpr int NSHAPE = 5, NSHAPE_CAND = 3;
constexpr int shape_value(int role, int c)
{
    return role == ICF_SHAPE_CLOSEST ? 4 << c : role == ICF_SHAPE_ORDER ? c : 64 << c;
}
constexpr int shape_ncand(int role) { return role == ICF_SHAPE_ORDER ? 2 : 3; }
constexpr int SHAPE_DEFAULT[NSHAPE] = {2, 2, 1, 2, 1};   // 256, 256, 8 lanes, 256, interleaved

struct icf_ctx {
    int id = 0;
    int device = 0;
    int ncu = 256;                 // compute units of the device
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipStream_t cur = nullptr;       // where LAUNCH issues: `stream`, or `stream_c` (lanes)
    // lanes across calls: lane 0 is `stream`, lane 1 is `stream_c` (contact detection); each
    // lane's log holds its calls that the other lane has not yet waited for
    hipStream_t stream_c = nullptr;
    hipEvent_t lane_ev[2][LANE_EVENTS] = {};
    unsigned lane_ev_next[2] = {0, 0};
    std::vector<LaneRec> lane_log[2];
    bool lane0_unlogged = false;   // lane-0 work issued since the last lane-0 log entry
    bool has_consts = false;
    bool consts_dirty = false;     // icf_set_constants changed hc since this context's last upload
    // bumped whenever memory the context owns is freed or moved, or a slot map or the
    // adjacency lists may no longer match what they were built from: a graph captured at an
    // older epoch may point at stale memory and is refused by icf_graph_launch
    unsigned long long epoch = 1;
    int precision = ICF_FP64;
    DevConsts hc;
    DevBuf tet_rhs, tet_tan, cz_rhs, cz_tan, ct, scan_tmp;
    DevBuf pat_a, pat_b;           // pattern build / broad phase / NL2set: 64-bit keys
    // the pattern's static part (elements and CZs): their distinct sorted keys, kept with a copy
    // of the node-id fields and altIds they came from; reused while that copy still matches
    struct PatCache {
        bool valid = false;
        int nNodes = 0, nElems = 0, el_stride = 0, nCZs = 0, cz_stride = 0, N = 0, cb = 0, nkeys = 0;
        DevBuf snap, keys;
    } patc;
    DevBuf bp_f, bp_i, bp_l;         // broad phase and NL2set scratch (doubles, ints, int64)
    DevBuf bp_c;                     // broad phase grid (cell counts, first sorted position per cell key)
    DevBuf bp_g;                     // broad phase pair grouping (counts and starts per tet, ranks)
    long long bp_cap = 0;            // pair keys the broad phase's buffer holds (grows with use)
    DevBuf tsum;                     // per-node tet sums of the fused call
    DevBuf ct_f;                     // impact forces (the tangent blocks are in ct)
    int tplane = 64;                 // plane stride of the staged tet records (tet_plane)
    // compact adjacent-face lists of the surface tets (k_adj_*), built from `ie` on first use and
    // dropped when a library write or free touches that allocation (or icf_pattern_changed)
    struct Adj {
        const int *ie = nullptr;
        int es = 0;
        uintptr_t lo = 0, hi = 0;    // the allocation holding ie
        DevBuf off, ids;
        bool valid = false;
    } adj;
    int *small = nullptr;          // 8 ints of scratch (max key, error flag, sums)
    unsigned char *noalloc = nullptr;   // this device's `scratch` array (device address)
    std::list<SlotMap> maps;       // std::list: pointers handed out by get_map stay valid
    int next_map_id = 0;
    std::string err;
    double last_ms = -1;
    bool blocking = true;          // icf_set_blocking
    bool capturing = false;        // between icf_graph_begin and icf_graph_end
    bool overlap = true;           // icf_set_overlap: independent kernels on more than one stream
    // autotuned launch shapes.  tune: the chunk shape (workgroup size, staging) of the BCSR-value
    // reductions, per role [accumulate map][FP64 staging]; shape: the workgroup size of phase 1,
    // the narrow phase and the impacts' phase 1, and the lanes per tuple of the closest face.
    // Per role: the candidate in use and, while tuning, the timings
    struct Tune {
        int ncand = NCAND;
        int sel = CAND_DEFAULT_FP32;   // candidate in use
        bool active = false;       // timing the candidates in turn
        bool fixed = false;        // chosen by the caller (icf_set_chunk_keys / icf_set_shape)
        int turn = 0;              // launches started while tuning
        int samples = 0;           // launches timed
        int cand = -1;             // candidate of the launch being timed, -1: none
        int bucket = -1;           // size class of the launches timed (the shape cache's key)
        float best[NCAND] = {};    // fastest launch per candidate (ms), 0: not timed
        hipEvent_t ev[2] = {nullptr, nullptr};
    } tune[2][2], shape[NSHAPE];
    int tune_rounds = 0;           // timed launches per candidate (icf_set_autotune)
    // icf_assemble_batched: the last offsets (host copy; device copy followed by 2 counts per
    // specimen) and the value and rhs maps whose units were last checked against them
    struct Batch {
        std::vector<int> off;      // 4 (n + 1): tets, CZs, rows, blocks
        int n = 0;
        DevBuf d;
        // the units and epoch whose block-diagonal layout was last checked against `off`
        std::vector<long long> checked;
        // set by icf_assemble_batched for its assembly: the fused value launch then also writes
        // the per-specimen CZ counts (spec_counted reports that it did)
        const int *cz_off = nullptr;
        int *spec_counts = nullptr;
        bool spec_counted = false;
    } batch;
    // per-kernel profiling (icf_set_profiling)
    bool profiling = false;
    struct Pending {
        const char *name;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::string> prof_names;
    std::vector<double> prof_ms;
    std::vector<long long> prof_n;
};

namespace {

int fail(icf_ctx *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->err = msg;
    return code;
}

#define HIPCHK(ctx, x)                                                                             \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess)                                                                      \
            return fail((ctx), (int)e_, std::string(#x) + ": " + hipGetErrorString(e_));           \
    } while (0)

int grow(icf_ctx *ctx, DevBuf &b, size_t bytes)
{
    if (b.bytes >= bytes) return 0;
    if (ctx->capturing) return fail(ctx, ICF_E_CAPTURE, "an internal buffer must grow: run the call once uncaptured");
    if (b.p) {
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream_c));
        HIPCHK(ctx, hipFree(b.p));
        b.p = nullptr;
        b.bytes = 0;
    }
    size_t want = bytes + bytes / 4 + 256;
    ctx->epoch++;
    HIPCHK(ctx, hipMalloc(&b.p, want));
    b.bytes = want;
    return 0;
}

void free_map(SlotMap &m)
{
    if (m.off) (void)hipFree(m.off);
    if (m.codes) (void)hipFree(m.codes);
    if (m.keys) (void)hipFree(m.keys);
    if (m.row_node) (void)hipFree(m.row_node);
    if (m.cz_begin) (void)hipFree(m.cz_begin);
    if (m.node_row) (void)hipFree(m.node_row);
    m.node_row = nullptr;
    m.merge_ok = false;
    if (m.chunks[0][0]) (void)hipFree(m.chunks[0][0]);   // the start of the one allocation
    for (int c = 0; c < NCAND; c++) m.chunks[c][0] = m.chunks[c][1] = nullptr;
    if (m.snap) (void)hipFree(m.snap);
    m.snap = nullptr;
    m.off = nullptr;
    m.codes = nullptr;
    m.keys = nullptr;
    m.row_node = m.cz_begin = nullptr;
}

// Maps whose key fields overlap [lo, hi): dropped when the memory goes away (erase), else
// marked suspect, to be compared against their snapshot before the next use.
bool map_keys_overlap(const SlotMap &m, uintptr_t lo, uintptr_t hi)
{
    for (const SrcKey &s : m.srcs) {
        uintptr_t a = (uintptr_t)(s.base + (size_t)s.stride * s.field0);
        uintptr_t b = (uintptr_t)(s.base + (size_t)s.stride * (s.field0 + s.nq));
        if (a < hi && lo < b) return true;
    }
    return false;
}

void invalidate_range(icf_ctx *ctx, uintptr_t lo, uintptr_t hi, bool erase = false)
{
    if (ctx->adj.valid && ctx->adj.lo < hi && lo < ctx->adj.hi) {
        ctx->adj.valid = false;
        ctx->epoch++;
    }
    for (auto it = ctx->maps.begin(); it != ctx->maps.end();) {
        const bool hit = map_keys_overlap(*it, lo, hi);
        if (hit) ctx->epoch++;
        if (hit && erase) {
            (void)hipStreamSynchronize(ctx->stream);
            free_map(*it);
            it = ctx->maps.erase(it);
        } else {
            if (hit) it->suspect = true;
            ++it;
        }
    }
}

// A write by the library into caller memory [p, p + bytes) (copies, kernels that write int
// fields in place).  Outside capture: slot maps and adjacency lists built from those bytes
// are checked again before their next use.  During capture such a write is refused: every
// replay would change what the maps were built from behind their back.
int note_write(icf_ctx *ctx, const void *p, size_t bytes, const char *what)
{
    const uintptr_t lo = (uintptr_t)p, hi = lo + bytes;
    if (!bytes) return 0;
    if (!ctx->capturing) {
        invalidate_range(ctx, lo, hi);
        return 0;
    }
    bool hit = ctx->adj.valid && ctx->adj.lo < hi && lo < ctx->adj.hi;
    for (const SlotMap &m : ctx->maps) hit = hit || map_keys_overlap(m, lo, hi);
    if (hit)
        return fail(ctx, ICF_E_CAPTURE, std::string(what) +
                                            " would write fields that a slot map or the adjacency lists were built "
                                            "from: it cannot be captured");
    return 0;
}

// The context's constants in g_c for the lifetime of this object: takes the device's lock,
// and when another context's constants are loaded, waits for all work on the device (its
// queued kernels read them) before loading ours.  During capture nothing is loaded: a replay
// loads them in icf_graph_launch.
struct ConstUse {
    std::unique_lock<std::mutex> lk;
    int err = 0;
    explicit ConstUse(icf_ctx *ctx)
    {
        if (!ctx->has_consts) {
            err = fail(ctx, ICF_E_NOCONST, "icf_set_constants has not been called");
            return;
        }
        if (ctx->capturing) return;
        lk = std::unique_lock<std::mutex>(g_dev_mu[ctx->device]);
        int &owner = g_const_owner[ctx->device];
        if (owner == ctx->id && !ctx->consts_dirty) return;
        hipError_t e = hipSuccess;
        if (owner != ctx->id && owner != 0) e = hipDeviceSynchronize();
        if (e == hipSuccess)
            e = hipMemcpyToSymbolAsync(HIP_SYMBOL(g_c), &ctx->hc, sizeof(DevConsts), 0, hipMemcpyHostToDevice,
                                       ctx->stream);
        if (e != hipSuccess) {
            owner = 0;
            err = fail(ctx, (int)e, std::string("loading the constants: ") + hipGetErrorString(e));
            return;
        }
        owner = ctx->id;
        ctx->consts_dirty = false;
    }
};

unsigned blocks_for(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

hipEvent_t prof_begin(icf_ctx *ctx)
{
    if (!ctx->profiling || ctx->capturing) return nullptr;
    hipEvent_t e = nullptr;
    if (!ctx->ev_pool.empty()) {
        e = ctx->ev_pool.back();
        ctx->ev_pool.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
        return nullptr;
    }
    (void)hipEventRecord(e, ctx->cur);
    return e;
}

void prof_end(icf_ctx *ctx, const char *name, hipEvent_t a)
{
    if (!a) return;
    hipEvent_t b = prof_begin(ctx);
    if (!b) {
        ctx->ev_pool.push_back(a);
        return;
    }
    ctx->pending.push_back({name, a, b});
}

// After the stream has drained: fold pending event pairs into the per-kernel totals.
void prof_flush(icf_ctx *ctx)
{
    for (auto &p : ctx->pending) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            size_t i = 0;
            while (i < ctx->prof_names.size() && ctx->prof_names[i] != p.name) i++;
            if (i == ctx->prof_names.size()) {
                ctx->prof_names.push_back(p.name);
                ctx->prof_ms.push_back(0);
                ctx->prof_n.push_back(0);
            }
            ctx->prof_ms[i] += ms;
            ctx->prof_n[i] += 1;
        }
        ctx->ev_pool.push_back(p.a);
        ctx->ev_pool.push_back(p.b);
    }
    ctx->pending.clear();
}

// Kernel launch on the context's stream, bracketed by events when profiling is on.
#define LAUNCH(ctx, name, kern, g, b, ...)                                   \
    do {                                                                     \
        hipEvent_t ea_ = prof_begin(ctx);                                    \
        hipLaunchKernelGGL(kern, g, b, 0, (ctx)->cur, __VA_ARGS__);          \
        prof_end(ctx, name, ea_);                                            \
    } while (0)

// Build (or fetch) the slot map for `srcs` with keys of `target`.
int get_map(icf_ctx *ctx, int target, int fixed_nkeys, const std::vector<SrcKey> &srcs, SlotMap **out)
{
    for (auto it = ctx->maps.begin(); it != ctx->maps.end(); ++it) {
        SlotMap &m = *it;
        if (m.target != target || m.fixed_nkeys != fixed_nkeys || !(m.srcs == srcs)) continue;
        if (m.suspect) {
            if (ctx->capturing)
                return fail(ctx, ICF_E_CAPTURE, "a slot map must be checked during capture: run the call once uncaptured");
            int *d_diff = ctx->small + 1;
            hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, ctx->stream, d_diff, 1, 0);
            long long off = 0;
            for (const SrcKey &k : srcs) {
                if (k.n == 0) continue;
                MapSrc ms{k.base, k.n, k.stride, k.field0, k.nq, k.kind};
                const long long tot = (long long)k.n * k.nq;
                hipLaunchKernelGGL(k_snap_cmp, dim3(blocks_for(tot, 256)), dim3(256), 0, ctx->stream, ms,
                                   m.snap + off, d_diff);
                off += tot;
            }
            int diff = 1;
            HIPCHK(ctx, hipGetLastError());
            HIPCHK(ctx, hipMemcpyAsync(&diff, d_diff, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
            if (diff) {
                free_map(m);
                ctx->maps.erase(it);
                ctx->epoch++;
                break;
            }
            m.suspect = false;
        }
        *out = &m;
        return 0;
    }
    if (ctx->capturing)
        return fail(ctx, ICF_E_CAPTURE, "a slot map would be built during capture: run the call once uncaptured");
    for (const SrcKey &s : srcs)
        if ((uint32_t)s.n > CODE_UNIT_MAX + 1u)
            return fail(ctx, ICF_E_CAPACITY, "more than 2^24 units in one pcsr array");

    SlotMap m;
    m.target = target;
    m.fixed_nkeys = fixed_nkeys;
    m.srcs = srcs;
    int *d_max = ctx->small, *d_err = ctx->small + 1;
    int nkeys = fixed_nkeys;
    if (nkeys < 0) {
        int init = -1;
        HIPCHK(ctx, hipMemcpyAsync(d_max, &init, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
        for (const SrcKey &s : srcs) {
            if (s.n == 0) continue;
            MapSrc ms{s.base, s.n, s.stride, s.field0, s.nq, s.kind};
            long long tot = (long long)s.n * s.nq;
            const unsigned nb = blocks_for(tot, 256);
            hipLaunchKernelGGL(k_map_maxkey, dim3(nb < 1024 ? nb : 1024), dim3(256), 0, ctx->stream, ms, d_max);
        }
        int mx = -1;
        HIPCHK(ctx, hipMemcpyAsync(&mx, d_max, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        nkeys = mx + 1;
    }
    m.nkeys = nkeys;
    HIPCHK(ctx, hipMalloc(&m.off, sizeof(int) * ((size_t)nkeys + 1)));
    int *d_cnt = nullptr;
    HIPCHK(ctx, hipMalloc(&d_cnt, sizeof(int) * ((size_t)nkeys + 1)));
    HIPCHK(ctx, hipMemsetAsync(d_cnt, 0, sizeof(int) * ((size_t)nkeys + 1), ctx->stream));
    HIPCHK(ctx, hipMemsetAsync(d_err, 0, sizeof(int), ctx->stream));
    for (const SrcKey &s : srcs) {
        if (s.n == 0) continue;
        MapSrc ms{s.base, s.n, s.stride, s.field0, s.nq, s.kind};
        long long tot = (long long)s.n * s.nq;
        hipLaunchKernelGGL(k_map_count, dim3(blocks_for(tot, 256)), dim3(256), 0, ctx->stream, ms, nkeys, d_cnt, d_err);
    }
    size_t tmp_bytes = 0;
    HIPCHK(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, d_cnt, m.off, nkeys + 1, ctx->stream));
    if (grow(ctx, ctx->scan_tmp, tmp_bytes)) {
        (void)hipFree(d_cnt);
        return (int)hipErrorOutOfMemory;
    }
    HIPCHK(ctx, hipcub::DeviceScan::ExclusiveSum(ctx->scan_tmp.p, tmp_bytes, d_cnt, m.off, nkeys + 1, ctx->stream));
    int total = 0, errflag = 0;
    HIPCHK(ctx, hipMemcpyAsync(&total, m.off + nkeys, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(&errflag, d_err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (errflag) {
        (void)hipFree(d_cnt);
        free_map(m);
        return fail(ctx, ICF_E_INDEX, "a slot/row/node index exceeds the declared size");
    }
    int *list = nullptr;           // accumulate mode: the keys with contributions
    int nlist = nkeys;
    if (fixed_nkeys < 0 && nkeys > 0) {
        // accumulate-mode maps list the keys that have contributions, so that their reduction
        // launches over those keys only
        int *flag = nullptr, *pos = nullptr;
        HIPCHK(ctx, hipMalloc(&flag, sizeof(int) * ((size_t)nkeys + 1)));
        HIPCHK(ctx, hipMalloc(&pos, sizeof(int) * ((size_t)nkeys + 1)));
        HIPCHK(ctx, hipMemsetAsync(flag + nkeys, 0, sizeof(int), ctx->stream));
        hipLaunchKernelGGL(k_map_flag, dim3(blocks_for(nkeys, 256)), dim3(256), 0, ctx->stream, d_cnt, nkeys, flag);
        HIPCHK(ctx, hipcub::DeviceScan::ExclusiveSum(ctx->scan_tmp.p, tmp_bytes, flag, pos, nkeys + 1, ctx->stream));
        int nt = 0;
        HIPCHK(ctx, hipMemcpyAsync(&nt, pos + nkeys, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        HIPCHK(ctx, hipMalloc(&list, sizeof(int) * ((size_t)nt + 1)));
        hipLaunchKernelGGL(k_map_compact, dim3(blocks_for(nkeys, 256)), dim3(256), 0, ctx->stream, flag, pos, nkeys,
                           list);
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        HIPCHK(ctx, hipFree(flag));
        HIPCHK(ctx, hipFree(pos));
        nlist = nt;
    }
    m.keys = list;                 // null when overwriting: every key, in order
    if (target == TGT_VALS && nlist > 0) {
        // chunks of consecutive list positions for the staged reduction of BCSR values, one
        // table per candidate key limit and staging type: at most `keys` keys and
        // flat_cap_of(FP32?, keys) contributions each (a single key with more contributions gets
        // a chunk of its own, reduced without staging)
        std::vector<int> h_off((size_t)nkeys + 1), h_list(list ? (size_t)nlist : 0);
        HIPCHK(ctx, hipMemcpyAsync(h_off.data(), m.off, sizeof(int) * ((size_t)nkeys + 1), hipMemcpyDeviceToHost,
                                   ctx->stream));
        if (list)
            HIPCHK(ctx, hipMemcpyAsync(h_list.data(), list, sizeof(int) * (size_t)nlist, hipMemcpyDeviceToHost,
                                       ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        // a chunk's codes are contiguous (keys in between, if any, have none): each entry is the
        // chunk's first list position and first code position; the last entry closes the table
        auto key_of = [&](int p) { return list ? h_list[p] : p; };
        std::vector<int2> ch;          // every table, one after the other, in one allocation
        size_t at[2 * NCAND];
        for (int kc = 0; kc < 2 * NCAND; kc++) {
            const int kind = kc & 1, keys = cand_keys(kc >> 1);
            const int cap = flat_cap_of(kind == 0, keys, cand_stage(kc >> 1));
            at[kc] = ch.size();
            ch.push_back(make_int2(0, h_off[key_of(0)]));
            int cnt = 0, nk = 0;
            for (int p = 0; p < nlist; p++) {
                const int k = key_of(p);
                const int c = h_off[k + 1] - h_off[k];
                if (nk > 0 && (nk == keys || cnt + c > cap)) {
                    ch.push_back(make_int2(p, h_off[k]));
                    cnt = 0;
                    nk = 0;
                }
                cnt += c;
                nk++;
            }
            ch.push_back(make_int2(nlist, h_off[key_of(nlist - 1) + 1]));
            m.nchunks[kc >> 1][kind] = (int)(ch.size() - at[kc]) - 1;
        }
        int2 *d = nullptr;
        HIPCHK(ctx, hipMalloc(&d, sizeof(int2) * ch.size()));
        for (int kc = 0; kc < 2 * NCAND; kc++) m.chunks[kc >> 1][kc & 1] = d + at[kc];
        HIPCHK(ctx, hipMemcpyAsync(d, ch.data(), sizeof(int2) * ch.size(), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    }
    m.nlaunch = nlist;
    m.ncodes = total;
    m.id = ++ctx->next_map_id;
    HIPCHK(ctx, hipMalloc(&m.codes, sizeof(uint32_t) * ((size_t)total + 1)));
    HIPCHK(ctx, hipMemsetAsync(d_cnt, 0, sizeof(int) * ((size_t)nkeys + 1), ctx->stream));
    for (const SrcKey &s : srcs) {
        if (s.n == 0) continue;
        MapSrc ms{s.base, s.n, s.stride, s.field0, s.nq, s.kind};
        long long tot = (long long)s.n * s.nq;
        hipLaunchKernelGGL(k_map_fill, dim3(blocks_for(tot, 256)), dim3(256), 0, ctx->stream, ms, nkeys, m.off, d_cnt,
                           m.codes);
    }
    if (nkeys > 0)
        hipLaunchKernelGGL(k_map_sort, dim3(blocks_for(nkeys, 256)), dim3(256), 0, ctx->stream, m.off, m.codes, nkeys);
    {
        long long snap_n = 0;
        for (const SrcKey &s : srcs) snap_n += (long long)s.n * s.nq;
        HIPCHK(ctx, hipMalloc(&m.snap, sizeof(int) * (size_t)(snap_n + 1)));
        long long off = 0;
        for (const SrcKey &s : srcs) {
            if (s.n == 0) continue;
            MapSrc ms{s.base, s.n, s.stride, s.field0, s.nq, s.kind};
            const long long tot = (long long)s.n * s.nq;
            hipLaunchKernelGGL(k_snap_copy, dim3(blocks_for(tot, 256)), dim3(256), 0, ctx->stream, ms, m.snap + off);
            off += tot;
        }
    }
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipFree(d_cnt));
    ctx->maps.push_back(m);
    *out = &ctx->maps.back();
    return 0;
}

// Whether the rows of rhs map mr can start from the per-node tet sums of force map mf (see
// k_rhs_check); decided once per pair of maps.
int get_tsum_link(icf_ctx *ctx, SlotMap *mr, const SlotMap *mf, const MapSrc &tet, bool *ok)
{
    *ok = false;
    if (mr->tsum_with != mf->id) {
        if (ctx->capturing)
            return fail(ctx, ICF_E_CAPTURE, "a slot map would be built during capture: run the call once uncaptured");
        if (mr->row_node) (void)hipFree(mr->row_node);
        if (mr->cz_begin) (void)hipFree(mr->cz_begin);
        if (mr->node_row) (void)hipFree(mr->node_row);
        mr->row_node = mr->cz_begin = mr->node_row = nullptr;
        mr->tsum_ok = mr->merge_ok = false;
        mr->tsum_with = mf->id;
        if (mr->nkeys > 0) {
            HIPCHK(ctx, hipMalloc(&mr->row_node, sizeof(int) * (size_t)mr->nkeys));
            HIPCHK(ctx, hipMalloc(&mr->cz_begin, sizeof(int) * (size_t)mr->nkeys));
            int *d_err = ctx->small + 1;
            hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, ctx->stream, d_err, 1, 0);
            hipLaunchKernelGGL(k_rhs_check, dim3(blocks_for(mr->nkeys, 256)), dim3(256), 0, ctx->stream, mr->off,
                               mr->codes, mr->nkeys, mf->off, mf->codes, mf->nkeys, tet, mr->row_node, mr->cz_begin,
                               d_err);
            int err = 1;
            HIPCHK(ctx, hipGetLastError());
            HIPCHK(ctx, hipMemcpyAsync(&err, d_err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
            mr->tsum_ok = err == 0;
            if (mr->tsum_ok && mf->nkeys > 0) {
                HIPCHK(ctx, hipMalloc(&mr->node_row, sizeof(int) * (size_t)mf->nkeys));
                hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, ctx->stream, d_err, 1, 0);
                hipLaunchKernelGGL(k_fill_int, dim3(blocks_for(mf->nkeys, 256)), dim3(256), 0, ctx->stream,
                                   mr->node_row, mf->nkeys, -1);
                hipLaunchKernelGGL(k_node_row, dim3(blocks_for(mr->nkeys, 256)), dim3(256), 0, ctx->stream,
                                   mr->row_node, mr->nkeys, mr->node_row, d_err);
                err = 1;
                HIPCHK(ctx, hipGetLastError());
                HIPCHK(ctx, hipMemcpyAsync(&err, d_err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
                HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
                mr->merge_ok = err == 0;
            }
        }
    }
    *ok = mr->tsum_ok;
    return 0;
}

int check_ctx(icf_ctx *ctx)
{
    if (!ctx) return ICF_E_INVALID;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    return 0;
}

struct Timer {
    icf_ctx *ctx;
    bool sync;
    explicit Timer(icf_ctx *c) : ctx(c), sync(c->blocking && !c->capturing)
    {
        ctx->last_ms = -1;
        if (sync) (void)hipEventRecord(ctx->ev0, ctx->stream);
    }
    int finish()
    {
        if (!sync) {
            HIPCHK(ctx, hipGetLastError());
            return 0;
        }
        HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
        HIPCHK(ctx, hipEventSynchronize(ctx->ev1));
        float ms = 0;
        HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ctx->last_ms = ms;
        prof_flush(ctx);
        HIPCHK(ctx, hipGetLastError());
        return 0;
    }
};

// ---- launch helpers ---------------------------------------------------------------------------

// ---- lanes across calls -----------------------------------------------------------------------
// When calls are enqueued without waiting (icf_set_blocking(0), outside capture, overlap on), the
// contact-detection work -- narrow phase, closest face, the per-impact phase of the collision
// response -- runs on lane 1 (`stream_c`) beside the assembly on lane 0 (`stream`).  Every call
// declares the device bytes it reads and writes; before issuing, it waits for the latest call of
// the other lane whose bytes conflict with its own (one of the two writes bytes the other
// touches).  Blocking calls finish before they return, and captured calls replay as one graph
// (whose independent branches the HIP runtime was measured to run one after the other), so
// both use lane 0 only; lane-0 work that was not logged is waited for in full by the next
// lane-1 call.

bool lanes_on(const icf_ctx *ctx) { return ctx->overlap && !ctx->profiling && !ctx->blocking && !ctx->capturing; }

hipStream_t lane_stream(const icf_ctx *ctx, int l) { return l ? ctx->stream_c : ctx->stream; }

// lane for a call that may run on lane 1
int contact_lane(const icf_ctx *ctx) { return lanes_on(ctx) ? 1 : 0; }

hipEvent_t next_lane_event(icf_ctx *ctx, int l) { return ctx->lane_ev[l][ctx->lane_ev_next[l]++ % LANE_EVENTS]; }

// `access` returns the call's Access; it is only evaluated when the other lane has logged calls
template <typename G> int lane_enter(icf_ctx *ctx, int l, G &&access)
{
    if (l == 1 && ctx->lane0_unlogged) {
        hipEvent_t ev = next_lane_event(ctx, 0);
        HIPCHK(ctx, hipEventRecord(ev, ctx->stream));
        HIPCHK(ctx, hipStreamWaitEvent(ctx->stream_c, ev, 0));
        ctx->lane_log[0].clear();
        ctx->lane0_unlogged = false;
    }
    std::vector<LaneRec> &other = ctx->lane_log[1 - l];
    if (!other.empty()) {
        auto &&a = access();
        for (int i = (int)other.size() - 1; i >= 0; i--) {
            if (!conflict(a, other[i].a)) continue;
            HIPCHK(ctx, hipStreamWaitEvent(lane_stream(ctx, l), other[i].ev, 0));
            // the other lane's calls up to i are now ordered before everything later on this lane
            other.erase(other.begin(), other.begin() + i + 1);
            break;
        }
    }
    ctx->cur = lane_stream(ctx, l);
    return 0;
}

template <typename G> int lane_exit(icf_ctx *ctx, int l, G &&access)
{
    ctx->cur = ctx->stream;
    if (l == 0 && !lanes_on(ctx)) {
        ctx->lane0_unlogged = true;
        return 0;
    }
    hipEvent_t ev = next_lane_event(ctx, l);
    HIPCHK(ctx, hipEventRecord(ev, lane_stream(ctx, l)));
    std::vector<LaneRec> &log = ctx->lane_log[l];
    if ((int)log.size() + 1 >= LANE_LOG) {
        // the latest event stands for every earlier call of the lane
        log.clear();
        log.push_back({Access::everything(), ev});
    } else {
        log.push_back({access(), ev});
    }
    return 0;
}

Access everything() { return Access::everything(); }

// A call that touches unknown memory, or returns host data: after all lane-1 work.
int lane_barrier(icf_ctx *ctx) { return lane_enter(ctx, 0, everything); }

// Both lanes idle, logs empty.
int sync_lanes(icf_ctx *ctx)
{
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream_c));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->lane_log[0].clear();
    ctx->lane_log[1].clear();
    ctx->lane0_unlogged = false;
    return 0;
}

// One call's lane: enters on construction (check `err`), done() logs it.  `make` builds the
// call's Access when a lane needs it.
template <typename F> struct OnLane {
    icf_ctx *ctx;
    int lane;
    F make;
    Access acc;
    bool have = false;
    int err = 0;
    OnLane(icf_ctx *c, int l, F f) : ctx(c), lane(l), make(f)
    {
        err = lane_enter(ctx, lane, [this]() -> const Access & { return get(); });
    }
    const Access &get()
    {
        if (!have) {
            acc = make();
            have = true;
        }
        return acc;
    }
    int done() { return lane_exit(ctx, lane, [this]() -> const Access & { return get(); }); }
    ~OnLane() { ctx->cur = ctx->stream; }
};

// reserve > n: only grow the staging buffers for `reserve` units
// Plane stride of the tet staging records for n tets.
int tet_plane(long long n) { return (int)((std::max(n, 1LL) + 63) / 64 * 64); }

// Staged tet records of more than this many bytes are written with nontemporal stores: they
// would not stay in the caches until phase 2 reads them anyway (measured: 3.5M tets faster
// with them, 384k tets faster without).
constexpr size_t STREAM_STAGING_BYTES = 128u << 20;
bool stream_tet_staging(const icf_ctx *ctx, int n)
{
    const size_t rec = (ctx->hc.iso ? TET_ISO_REC : TET_GEN_REC) * (ctx->precision == ICF_MIXED ? 4 : 8);
    return (size_t)tet_plane(n) * (rec + sizeof(double) * TET_RHS_REC) > STREAM_STAGING_BYTES;
}

int stage_tets(icf_ctx *ctx, const int *ie, const double *dn, int n, int es, int ns, int reserve = 0)
{
    const bool mixed = ctx->precision == ICF_MIXED, iso = ctx->hc.iso != 0;
    const size_t rec = iso ? TET_ISO_REC : TET_GEN_REC;
    const size_t cap = (size_t)tet_plane(std::max(n, reserve));
    if (int r = grow(ctx, ctx->tet_rhs, sizeof(double) * TET_RHS_REC * cap)) return r;
    if (int r = grow(ctx, ctx->tet_tan, (mixed ? 4 : 8) * rec * cap)) return r;
    if (n == 0) return 0;
    const int tp = ctx->tplane = tet_plane(n);
    dim3 g(blocks_for(n, 256)), b(256);
    double *sr = (double *)ctx->tet_rhs.p;
    void *st = ctx->tet_tan.p;
    const bool so = stream_tet_staging(ctx, n);
#define ICF_T1(MX, IS, NAME)                                                                                   \
    do {                                                                                                      \
        if (so) LAUNCH(ctx, NAME, (k_tet_phase1<MX, IS, true>), g, b, ie, dn, n, tp, es, ns, sr, st);         \
        else LAUNCH(ctx, NAME, (k_tet_phase1<MX, IS, false>), g, b, ie, dn, n, tp, es, ns, sr, st);           \
    } while (0)
    if (mixed && iso) ICF_T1(true, true, "k_tet_phase1<mixed,iso>");
    else if (mixed) ICF_T1(true, false, "k_tet_phase1<mixed,gen>");
    else if (iso) ICF_T1(false, true, "k_tet_phase1<fp64,iso>");
    else ICF_T1(false, false, "k_tet_phase1<fp64,gen>");
#undef ICF_T1
    return 0;
}

int stage_czs(icf_ctx *ctx, double *dcz, int *icz, const double *dn, int n, int cs, int ns, int reserve = 0)
{
    const bool mixed = ctx->precision == ICF_MIXED;
    const size_t cap = (size_t)std::max(std::max(n, reserve), 1);
    if (int r = grow(ctx, ctx->cz_rhs, sizeof(double) * CZ_RHS_REC * cap)) return r;
    if (int r = grow(ctx, ctx->cz_tan, (mixed ? 4 : 8) * CZ_TAN_REC * cap)) return r;
    if (n == 0) return 0;
    const long long waves = (n + CZ_PER_WAVE - 1) / CZ_PER_WAVE;   // three lanes per CZ
    dim3 g(blocks_for(waves * 64, 256)), b(256);
    if (mixed)
        LAUNCH(ctx, "k_cz_phase1<mixed>", (k_cz_phase1<true>), g, b, dcz, icz, dn, n, cs, ns, (double *)ctx->cz_rhs.p,
               ctx->cz_tan.p);
    else
        LAUNCH(ctx, "k_cz_phase1<fp64>", (k_cz_phase1<false>), g, b, dcz, icz, dn, n, cs, ns, (double *)ctx->cz_rhs.p,
               ctx->cz_tan.p);
    return 0;
}

// ---- autotuner of the launch shapes -----------------------------------------------------------
// While a role is being tuned, its launches cycle through the role's candidates, each timed
// with a pair of events around the launch; the next launch of the role reads the previous
// timing (waiting for it if it has not finished), and after tune_rounds timings per candidate
// the fastest launch decides.  Captured calls are not timed; they use the current choice.
// Roles: the value reductions' chunk shape per (accumulate map, FP64 staging) (eight
// candidates), and the shape roles (three candidates each).  A finished tuning is also kept in
// a cache held by the library, per device, role, precision and size class (the launch's units,
// to the next power of two): contexts whose role is neither being tuned nor set by the caller
// take the cached choice for launches of that size class.

struct ShapeCache {
    std::mutex mu;
    std::map<std::array<int, 4>, int> choice;   // (device, role id, precision, size class) -> candidate
};
ShapeCache g_shape_cache;

int size_class(long long n)
{
    int b = 0;
    while (b < 62 && (1LL << b) < n) b++;
    return b;
}

void tune_fold(icf_ctx *ctx, icf_ctx::Tune &t, int role_id)
{
    if (t.cand >= 0) {
        float ms = 0;
        if (hipEventSynchronize(t.ev[1]) == hipSuccess && hipEventElapsedTime(&ms, t.ev[0], t.ev[1]) == hipSuccess &&
            ms > 0) {
            if (t.best[t.cand] == 0 || ms < t.best[t.cand]) t.best[t.cand] = ms;
            t.samples++;
        }
        t.cand = -1;
    }
    if (t.active && (t.samples >= ctx->tune_rounds * t.ncand || t.turn >= 4 * ctx->tune_rounds * t.ncand)) {
        t.active = false;
        for (int c = 0; c < t.ncand; c++)
            if (t.best[c] > 0 && (t.best[t.sel] == 0 || t.best[c] < t.best[t.sel])) t.sel = c;
        if (t.bucket >= 0) {
            std::lock_guard<std::mutex> lk(g_shape_cache.mu);
            g_shape_cache.choice[{ctx->device, role_id, ctx->precision, t.bucket}] = t.sel;
        }
    }
}

// The candidate for the next launch of a role with `units` units; starts its timing if tuning
// (not when `peek`).
int tune_begin(icf_ctx *ctx, icf_ctx::Tune &t, int role_id, long long units, bool peek = false)
{
    const int bucket = size_class(units);
    if (!t.active || ctx->capturing || peek) {
        if (!t.active && !t.fixed) {
            std::lock_guard<std::mutex> lk(g_shape_cache.mu);
            auto it = g_shape_cache.choice.find({ctx->device, role_id, ctx->precision, bucket});
            if (it != g_shape_cache.choice.end()) return it->second;
        }
        return t.sel;
    }
    tune_fold(ctx, t, role_id);
    if (!t.active) return t.sel;
    t.bucket = bucket;
    t.cand = t.turn++ % t.ncand;
    (void)hipEventRecord(t.ev[0], ctx->cur);
    return t.cand;
}

void tune_end(icf_ctx *ctx, icf_ctx::Tune &t)
{
    if (t.cand >= 0 && !ctx->capturing) (void)hipEventRecord(t.ev[1], ctx->cur);
}

// value-reduction roles: ids 0-3
int tune_begin(icf_ctx *ctx, bool acc, bool fp64, long long keys)
{
    return tune_begin(ctx, ctx->tune[acc][fp64], 2 * acc + fp64, keys);
}
void tune_end(icf_ctx *ctx, bool acc, bool fp64) { tune_end(ctx, ctx->tune[acc][fp64]); }

// shape roles: ids 4-7; returns the role's value (workgroup size or lanes per tuple)
int shape_begin(icf_ctx *ctx, int role, long long units, bool peek = false)
{
    return shape_value(role, tune_begin(ctx, ctx->shape[role], 4 + role, units, peek));
}
void shape_end(icf_ctx *ctx, int role) { tune_end(ctx, ctx->shape[role]); }

// ---- fused launches (icf_fused.inc.h) ----------------------------------------------------------

int round8(long long n) { return (int)((n + 7) / 8 * 8); }

// Phase 1 of the fused call as one launch; the staging buffers have already grown.
void phase1_fused(icf_ctx *ctx, const int *ie, const double *dn, int nt, int es, int ns, double *dcz, int *icz, int nc,
                  int cs, int *counts)
{
    const bool mixed = ctx->precision == ICF_MIXED, iso = ctx->hc.iso != 0;
    const int B = shape_begin(ctx, ICF_SHAPE_PHASE1, (long long)nt + nc);
    const int nbt = nt ? round8(blocks_for(nt, B)) : 0;
    const int nbc = nc ? (int)blocks_for((long long)(nc + CZ_PER_WAVE - 1) / CZ_PER_WAVE * 64, B) : 0;
    const dim3 g(std::max(interleave2_grid(nbt, nbc), 1)), b(B);
    double *tr = (double *)ctx->tet_rhs.p, *cr = (double *)ctx->cz_rhs.p;
    void *tt = ctx->tet_tan.p, *ct = ctx->cz_tan.p;
    const int tp = ctx->tplane = tet_plane(nt);
    const bool so = stream_tet_staging(ctx, nt);
#define ICF_P1(MX, IS, NAME)                                                                                      \
    do {                                                                                                          \
        if (so)                                                                                                   \
            LAUNCH(ctx, NAME, (k_phase1_fused<MX, IS, true>), g, b, nbt, nbc, ie, dn, nt, tp, es, ns, tr, tt, dcz, \
                   icz, nc, cs, cr, ct, counts);                                                                  \
        else                                                                                                      \
            LAUNCH(ctx, NAME, (k_phase1_fused<MX, IS, false>), g, b, nbt, nbc, ie, dn, nt, tp, es, ns, tr, tt, dcz, \
                   icz, nc, cs, cr, ct, counts);                                                                  \
    } while (0)
    if (mixed && iso) ICF_P1(true, true, "k_phase1<mixed,iso>");
    else if (mixed) ICF_P1(true, false, "k_phase1<mixed,gen>");
    else if (iso) ICF_P1(false, true, "k_phase1<fp64,iso>");
    else ICF_P1(false, false, "k_phase1<fp64,gen>");
#undef ICF_P1
    shape_end(ctx, ICF_SHAPE_PHASE1);
}

// Calls f(integral_constant<cand_keys(c)>, integral_constant<cand_stage(c)>).
template <typename F>
void with_cand(int c, F &&f)
{
    static_assert(NCAND == 8, "one case per candidate");
    using std::integral_constant;
    switch (c) {
    case 0: f(integral_constant<int, cand_keys(0)>(), integral_constant<int, cand_stage(0)>()); break;
    case 1: f(integral_constant<int, cand_keys(1)>(), integral_constant<int, cand_stage(1)>()); break;
    case 2: f(integral_constant<int, cand_keys(2)>(), integral_constant<int, cand_stage(2)>()); break;
    case 3: f(integral_constant<int, cand_keys(3)>(), integral_constant<int, cand_stage(3)>()); break;
    case 4: f(integral_constant<int, cand_keys(4)>(), integral_constant<int, cand_stage(4)>()); break;
    case 5: f(integral_constant<int, cand_keys(5)>(), integral_constant<int, cand_stage(5)>()); break;
    case 6: f(integral_constant<int, cand_keys(6)>(), integral_constant<int, cand_stage(6)>()); break;
    default: f(integral_constant<int, cand_keys(7)>(), integral_constant<int, cand_stage(7)>()); break;
    }
}

// BCSR values (overwrite) and node forces with their tet sums as one launch.  False when the
// value map is an accumulate map: the caller launches them apart.
bool vals_force_fused(icf_ctx *ctx, const SlotMap *mv, double *A, const SlotMap *mf, double *F, int ns,
                      const Stage &st)
{
    if (mv->keys) return false;
    if (mv->nlaunch == 0 && mf->nlaunch == 0) return true;
    const bool mixed = ctx->precision == ICF_MIXED, iso = ctx->hc.iso != 0;
    const int k = mixed ? 0 : 1;
    double *ts = (double *)ctx->tsum.p;
    const int c = tune_begin(ctx, false, !mixed, mv->nlaunch);
    const int ilv = shape_begin(ctx, ICF_SHAPE_ORDER, mv->nlaunch);
    with_cand(c, [&](auto kc, auto sc) {
        constexpr int K = decltype(kc)::value, S = decltype(sc)::value;
        const int nch = mv->nlaunch ? mv->nchunks[c][k] : 0;
        const int nbf = mf->nlaunch ? round8(blocks_for(mf->nlaunch, K)) : 0;
        const dim3 g(interleave2_grid(nbf, nch)), b(K);   // not empty: at least one map has keys
#define ICF_VF(MX, IS, L, NAME)                                                                                   \
    LAUNCH(ctx, NAME, (k_vals_force_fused<MX, IS, L, K, S>), g, b, nbf, mv->off, mv->codes, mv->chunks[c][k], nch, A, \
           mf->off, mf->codes, mf->keys, mf->nlaunch, F, ns, st, ts, ilv)
        if (mixed && iso) ICF_VF(true, true, float, "k_vals_force<mixed,iso>");
        else if (mixed) ICF_VF(true, false, float, "k_vals_force<mixed,gen>");
        else if (iso) ICF_VF(false, true, double, "k_vals_force<fp64,iso>");
        else ICF_VF(false, false, double, "k_vals_force<fp64,gen>");
#undef ICF_VF
    });
    shape_end(ctx, ICF_SHAPE_ORDER);
    tune_end(ctx, false, !mixed);
    return true;
}

// The second half of the fused call as one launch: node forces with the rhs rows, the CZ
// counters and the BCSR values (overwrite).  False unless the rhs map's rows each have their
// tet node (merge_ok) and the value map is an overwrite map.
bool vals_force_rhs_fused(icf_ctx *ctx, const SlotMap *mv, double *A, const SlotMap *mf, double *F, int ns,
                          const SlotMap *mr, double *b, const int *dmg, const int *fld, int nCZs, int *counts,
                          const Stage &st)
{
    if (mv->keys || !mr->merge_ok || !mr->node_row) return false;
    if (mv->nlaunch == 0 && mf->nlaunch == 0 && nCZs == 0) return true;
    const bool mixed = ctx->precision == ICF_MIXED, iso = ctx->hc.iso != 0;
    const int k = mixed ? 0 : 1;
    const int c = tune_begin(ctx, false, !mixed, mv->nlaunch);
    const int ilv = shape_begin(ctx, ICF_SHAPE_ORDER, mv->nlaunch);
    with_cand(c, [&](auto kc, auto sc) {
        constexpr int K = decltype(kc)::value, S = decltype(sc)::value;
        const int nch = mv->nlaunch ? mv->nchunks[c][k] : 0;
        const int nbf = mf->nlaunch ? round8(blocks_for(mf->nlaunch, K)) : 0;
        // the CZ counters' workgroups: per specimen in a batched call (one or more specimens
        // each), else strided over all CZs
        const auto &bt = ctx->batch;
        const bool per_spec = bt.cz_off && nCZs;
        const int nbc = per_spec ? round8(std::min(bt.n, 64))
                                 : nCZs ? round8(std::min<long long>(blocks_for(nCZs, K), 64)) : 0;
        const dim3 g(nbc + interleave2_grid(nbf, nch)), bl(K);   // not empty: something has keys or CZs
#define ICF_VFR(MX, IS, L, NAME)                                                                                   \
    LAUNCH(ctx, NAME, (k_vals_force_rhs_fused<MX, IS, L, K, S>), g, bl, nbf, nbc, mv->off, mv->codes, mv->chunks[c][k], \
           nch, A, mf->off, mf->codes, mf->keys, mf->nlaunch, F, ns, mr->node_row, mr->off, mr->codes, mr->cz_begin, \
           b, dmg, fld, nCZs, counts, per_spec ? bt.cz_off : nullptr, bt.n, bt.spec_counts, st, ilv)
        if (mixed && iso) ICF_VFR(true, true, float, "k_vals_force_rhs<mixed,iso>");
        else if (mixed) ICF_VFR(true, false, float, "k_vals_force_rhs<mixed,gen>");
        else if (iso) ICF_VFR(false, true, double, "k_vals_force_rhs<fp64,iso>");
        else ICF_VFR(false, false, double, "k_vals_force_rhs<fp64,gen>");
#undef ICF_VFR
        ctx->batch.spec_counted = per_spec;
    });
    shape_end(ctx, ICF_SHAPE_ORDER);
    tune_end(ctx, false, !mixed);
    return true;
}

// The collision response's sums into node forces, BCSR blocks and rhs rows as one launch.
// False when the value map is an overwrite map.
bool contact_sums_fused(icf_ctx *ctx, const SlotMap *mf, double *F, int ns, const SlotMap *mv, double *A,
                        const SlotMap *mr, double *b, const Stage &st)
{
    if (!mv->keys) return false;
    if (mf->nlaunch == 0 && mv->nlaunch == 0 && mr->nlaunch == 0) return true;
    const bool mixed = ctx->precision == ICF_MIXED;
    const int k = mixed ? 0 : 1;
    const int c = tune_begin(ctx, true, !mixed, mv->nlaunch);
    with_cand(c, [&](auto kc, auto sc) {
        constexpr int K = decltype(kc)::value, S = decltype(sc)::value;
        const int nch = mv->nlaunch ? mv->nchunks[c][k] : 0;
        const int nb0 = mf->nlaunch ? round8(blocks_for(mf->nlaunch, K)) : 0, nb1 = round8(nch);
        const int nb2 = mr->nlaunch ? (int)blocks_for(mr->nlaunch, K) : 0;
        const dim3 g(nb0 + nb1 + nb2), bl(K);
#define ICF_CS(MX, L, NAME)                                                                                       \
    LAUNCH(ctx, NAME, (k_contact_sums<MX, L, K, S>), g, bl, nb0, nb1, mf->off, mf->codes, mf->keys, mf->nlaunch, F, ns, \
           mv->off, mv->codes, mv->keys, mv->chunks[c][k], nch, A, mr->off, mr->codes, mr->keys, mr->nlaunch, b, st)
        if (mixed) ICF_CS(true, float, "k_contact_sums<mixed>");
        else ICF_CS(false, double, "k_contact_sums<fp64>");
#undef ICF_CS
    });
    tune_end(ctx, true, !mixed);
    return true;
}

Stage make_stage(icf_ctx *ctx, double h)
{
    Stage sv;
    sv.tet_rhs = (const double *)ctx->tet_rhs.p;
    sv.tet_tan = ctx->tet_tan.p;
    sv.tplane = ctx->tplane;
    sv.cz_rhs = (const double *)ctx->cz_rhs.p;
    sv.cz_tan = ctx->cz_tan.p;
    sv.ct_tan = ctx->ct.p;
    sv.ct_f = (const double *)ctx->ct_f.p;
    const DevConsts &c = ctx->hc;
    sv.stiff = 1.0 + c.dampK * c.ngamma / (h * c.nbeta);
    sv.mcoef = (1.0 / (h * h) + c.dampM * c.ngamma / h) / c.nbeta;
    return sv;
}

template <bool MX, bool IS, bool ACC, typename L>
void vals_kernel(icf_ctx *ctx, const char *name, const SlotMap *m, double *A, const Stage &sv)
{
    const int k = sizeof(L) == 4 ? 0 : 1;
    if (m->nlaunch == 0) return;   // then no chunks for any key limit
    const int c = tune_begin(ctx, ACC, k == 1, m->nlaunch);
    with_cand(c, [&](auto kc, auto sc) {
        constexpr int K = decltype(kc)::value, S = decltype(sc)::value;
        const int n = m->nchunks[c][k];
        LAUNCH(ctx, name, (k_reduce_vals<MX, IS, ACC, L, K, S>), dim3(n), dim3(K), m->off, m->codes, m->keys,
               m->chunks[c][k], n, A, sv);
    });
    tune_end(ctx, ACC, k == 1);
}

// Phase 2 of BCSR values for map m.  accumulate: add into A over the keys with contributions
// (single-kernel entry points); otherwise overwrite every block.  contact: the map's units
// are impacts, whose FP64 records are staged as FP64 in both modes.
void launch_vals(icf_ctx *ctx, const SlotMap *m, double *A, int accumulate, const Stage &sv, bool contact = false)
{
    if (m->nlaunch == 0) return;
    const bool mixed = ctx->precision == ICF_MIXED, iso = ctx->hc.iso != 0;
    if (contact) {
        // impact records are FP64; in mixed mode their blocks are staged (and added) as FP32
        // values, like the other tangent blocks of that mode
        if (mixed) vals_kernel<true, false, true, float>(ctx, "k_reduce_vals<contact,mixed,acc>", m, A, sv);
        else vals_kernel<false, false, true, double>(ctx, "k_reduce_vals<contact,fp64,acc>", m, A, sv);
        return;
    }
#define ICF_VALS(MX, IS, L, NAME)                                                                              \
    do {                                                                                                       \
        if (accumulate)                                                                                        \
            vals_kernel<MX, IS, true, L>(ctx, "k_reduce_vals<" NAME ",acc>", m, A, sv);                        \
        else                                                                                                   \
            vals_kernel<MX, IS, false, L>(ctx, "k_reduce_vals<" NAME ">", m, A, sv);                           \
    } while (0)
    if (mixed && iso) ICF_VALS(true, true, float, "mixed,iso");
    else if (mixed) ICF_VALS(true, false, float, "mixed,gen");
    else if (iso) ICF_VALS(false, true, double, "fp64,iso");
    else ICF_VALS(false, false, double, "fp64,gen");
#undef ICF_VALS
}

void launch_vec3(icf_ctx *ctx, const SlotMap *m, double *out, int comp_stride, int elem_stride, int accumulate,
                 const Stage &sv)
{
    const int *keys = m->keys;
    const int n = m->nlaunch;
    if (n == 0) return;
    dim3 g(blocks_for(n, 256)), b(256);
    const bool force = elem_stride == 1;
    if (accumulate)
        LAUNCH(ctx, force ? "k_reduce_vec3<force,acc>" : "k_reduce_vec3<rhs,acc>", k_reduce_vec3<true>, g, b, m->off,
               m->codes, keys, n, out, comp_stride, elem_stride, sv);
    else
        LAUNCH(ctx, force ? "k_reduce_vec3<force>" : "k_reduce_vec3<rhs>", k_reduce_vec3<false>, g, b, m->off,
               m->codes, keys, n, out, comp_stride, elem_stride, sv);
}

bool bad_units(int n, int stride) { return n < 0 || stride < 0 || (n > 0 && stride < n); }

}  // namespace

// =================================================================================================
// C ABI
// =================================================================================================
extern "C" {

int icf_create(int device, icf_ctx **out)
{
    if (!out || device < 0 || device >= 64) return ICF_E_INVALID;
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess) return (int)e;
    if (device >= ndev) return ICF_E_INVALID;
    icf_ctx *ctx = new icf_ctx();
    for (auto &row : ctx->tune) row[1].sel = CAND_DEFAULT_FP64;   // FP64 staging roles
    for (int r = 0; r < NSHAPE; r++) {
        ctx->shape[r].ncand = shape_ncand(r);
        ctx->shape[r].sel = SHAPE_DEFAULT[r];
    }
    ctx->id = g_next_ctx_id++;
    ctx->device = device;
    memset(&ctx->hc, 0, sizeof(ctx->hc));
    e = hipSetDevice(device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ctx->ncu, hipDeviceAttributeMultiprocessorCount, device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream_c, hipStreamNonBlocking);
    for (int l = 0; l < 2; l++)
        for (int k = 0; k < LANE_EVENTS && e == hipSuccess; k++)
            e = hipEventCreateWithFlags(&ctx->lane_ev[l][k], hipEventDisableTiming);


    ctx->cur = ctx->stream;
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev0);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev1);
    if (e == hipSuccess) e = hipMalloc(&ctx->small, 8 * sizeof(int));
    if (e == hipSuccess) e = scratch_for_device(device, (void **)&ctx->noalloc);
    if (e != hipSuccess) {
        delete ctx;
        return (int)e;
    }
    *out = ctx;
    return 0;
}

int icf_destroy(icf_ctx *ctx)
{
    if (!ctx) return ICF_E_INVALID;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream_c) (void)hipStreamSynchronize(ctx->stream_c);
    (void)hipStreamSynchronize(ctx->stream);
    for (SlotMap &m : ctx->maps) free_map(m);
    DevBuf *bufs[] = {&ctx->tet_rhs, &ctx->tet_tan, &ctx->cz_rhs, &ctx->cz_tan, &ctx->ct,   &ctx->scan_tmp,
                      &ctx->pat_a,   &ctx->pat_b,   &ctx->bp_f,   &ctx->bp_i, &ctx->bp_l, &ctx->bp_c,
                      &ctx->tsum,    &ctx->ct_f,    &ctx->adj.off, &ctx->adj.ids, &ctx->batch.d,
                      &ctx->patc.snap, &ctx->patc.keys, &ctx->bp_g};
    for (DevBuf *b : bufs)
        if (b->p) (void)hipFree(b->p);
    if (ctx->small) (void)hipFree(ctx->small);
    (void)hipEventDestroy(ctx->ev0);
    (void)hipEventDestroy(ctx->ev1);
    for (auto &p : ctx->pending) {
        ctx->ev_pool.push_back(p.a);
        ctx->ev_pool.push_back(p.b);
    }
    for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
    for (auto &row : ctx->tune)
        for (auto &t : row)
            for (hipEvent_t e : t.ev)
                if (e) (void)hipEventDestroy(e);
    for (auto &t : ctx->shape)
        for (hipEvent_t e : t.ev)
            if (e) (void)hipEventDestroy(e);
    for (int l = 0; l < 2; l++)
        for (int k = 0; k < LANE_EVENTS; k++)
            if (ctx->lane_ev[l][k]) (void)hipEventDestroy(ctx->lane_ev[l][k]);
    if (ctx->stream_c) (void)hipStreamDestroy(ctx->stream_c);
    (void)hipStreamDestroy(ctx->stream);
    {
        std::lock_guard<std::mutex> lk(g_dev_mu[ctx->device]);
        if (g_const_owner[ctx->device] == ctx->id) g_const_owner[ctx->device] = 0;
    }
    delete ctx;
    return 0;
}

const char *icf_last_error_string(const icf_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int icf_set_constants(icf_ctx *ctx, const icf_consts *c)
{
    if (!ctx || !c) return ICF_E_INVALID;
    // E symmetric
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++)
            if (c->E[i * 6 + j] != c->E[j * 6 + i]) return fail(ctx, ICF_E_CONSTANTS, "E is not symmetric");
    // B must be [s (x) I3, -s (x) I3] built from sf (ModelPrms.k_Bmatrix)
    for (int gp = 0; gp < 3; gp++)
        for (int p = 0; p < 3; p++)
            for (int k = 0; k < 18; k++) {
                int node = k / 3;
                double want = 0;
                if (k % 3 == p) want = node < 3 ? c->sf[node * 3 + gp] : -c->sf[(node - 3) * 3 + gp];
                if (c->B[gp * 54 + p * 18 + k] != want)
                    return fail(ctx, ICF_E_CONSTANTS, "CZ B table is not the shape-function Kronecker form");
            }
    DevConsts &d = ctx->hc;
    memset(&d, 0, sizeof(d));
    memcpy(d.E, c->E, sizeof(d.E));
    memcpy(d.M, c->M, sizeof(d.M));
    d.nbeta = c->NewmarkBeta; d.ngamma = c->NewmarkGamma;
    d.dampM = c->dampingMass; d.dampK = c->dampingStiffness;
    d.rho = c->rho; d.grav = c->gravity;
    d.G_fn = c->G_fn; d.G_ft = c->G_ft; d.f_tn = c->f_tn; d.f_tt = c->f_tt;
    d.alpha = c->alpha; d.beta = c->beta; d.rn = c->rn; d.rt = c->rt;
    d.p_m = c->p_m; d.p_n = c->p_n; d.deln = c->deln; d.delt = c->delt;
    d.pMtn = c->pMtn; d.pMnt = c->pMnt; d.gam_n = c->gam_n; d.gam_t = c->gam_t;
    memcpy(d.sf, c->sf, sizeof(d.sf));
    // isotropic split: E = lam 1 1^T (normal block) + diag(2mu,2mu,2mu,mu,mu,mu)
    {
        double lam = c->E[1], mu = c->E[3 * 6 + 3], dg = c->E[0];
        bool iso = mu > 0;
        for (int i = 0; i < 6 && iso; i++)
            for (int j = 0; j < 6 && iso; j++) {
                double v = c->E[i * 6 + j];
                if (i < 3 && j < 3) iso = (i == j) ? (v == dg) : (v == lam);
                else if (i == j) iso = (v == mu);
                else iso = (v == 0);
            }
        if (iso) iso = fabs(dg - (lam + 2 * mu)) <= 1e-12 * fabs(dg);
        // the isotropic path also takes the mass as m_ij I3 per node pair (ModelPrms.cs:239-248
        // builds it so); any other M takes the general path
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) {
                const double mij = c->M[(3 * i) * 12 + 3 * j];
                d.m4[4 * i + j] = mij;
                d.m4f[4 * i + j] = (float)mij;
                for (int k = 0; k < 3 && iso; k++)
                    for (int l = 0; l < 3 && iso; l++)
                        iso = c->M[(3 * i + k) * 12 + 3 * j + l] == (k == l ? mij : 0.0);
            }
        d.iso = iso ? 1 : 0;
        d.lam = lam;
        d.mu = mu;
        d.lamf = (float)lam;
        d.muf = (float)mu;
    }
    d.k_tn = c->gam_n / c->deln;
    d.k_tt = c->gam_t / c->delt;
    d.k_dnn = c->gam_n / (c->deln * c->deln);
    d.k_dtt = c->gam_t / (c->delt * c->delt);
    d.k_dnt = c->gam_n * c->gam_t / (c->deln * c->delt);
    d.pm_alpha = c->p_m / c->alpha;
    d.pn_beta = c->p_n / c->beta;
    d.thr_n = c->f_tn * 0.05;
    d.thr_t = c->f_tt * 0.05;
    d.lim_n = c->deln * c->rn / 5;
    d.lim_t = c->delt * c->rt / 5;
    for (int e = 0; e < 3; e++) {
        d.bn0[e] = pow(c->p_m / c->alpha, c->p_m - e);
        d.bt0[e] = pow(c->p_n / c->beta, c->p_n - e);
    }
    for (int i = 0; i < 36; i++) d.Ef[i] = (float)d.E[i];
    for (int i = 0; i < 144; i++) d.Mf[i] = (float)d.M[i];
    ctx->has_consts = true;
    ctx->consts_dirty = true;
    if (int r = check_ctx(ctx)) return r;
    ConstUse cu(ctx);
    return cu.err;
}

int icf_set_precision(icf_ctx *ctx, int mode)
{
    if (!ctx || (mode != ICF_FP64 && mode != ICF_MIXED)) return ICF_E_INVALID;
    ctx->precision = mode;
    return 0;
}

int icf_get_precision(const icf_ctx *ctx) { return ctx ? ctx->precision : ICF_E_INVALID; }

int icf_alloc(icf_ctx *ctx, size_t bytes, void **dptr)
{
    if (!ctx || !dptr) return ICF_E_INVALID;
    if (int r = check_ctx(ctx)) return r;
    *dptr = nullptr;
    HIPCHK(ctx, hipMalloc(dptr, bytes ? bytes : 1));
    return 0;
}

int icf_free(icf_ctx *ctx, void *dptr)
{
    if (!ctx) return ICF_E_INVALID;
    if (!dptr) return 0;
    if (ctx->capturing) return fail(ctx, ICF_E_CAPTURE, "icf_free during capture");
    if (int r = check_ctx(ctx)) return r;
    if (int r = sync_lanes(ctx)) return r;
    size_t sz = 0;
    void *basep = nullptr;
    if (hipMemGetAddressRange(&basep, &sz, dptr) == hipSuccess && sz)
        invalidate_range(ctx, (uintptr_t)basep, (uintptr_t)basep + sz, true);
    else
        invalidate_range(ctx, (uintptr_t)dptr, (uintptr_t)dptr + 1, true);
    HIPCHK(ctx, hipFree(dptr));
    return 0;
}

int icf_h2d(icf_ctx *ctx, void *dst, size_t dst_off, const void *src, size_t bytes)
{
    if (!ctx || (!dst && bytes) || (!src && bytes)) return ICF_E_INVALID;
    if (ctx->capturing) return fail(ctx, ICF_E_CAPTURE, "host copies cannot be captured");
    if (int r = check_ctx(ctx)) return r;
    if (!bytes) return 0;
    char *d = (char *)dst + dst_off;
    invalidate_range(ctx, (uintptr_t)d, (uintptr_t)d + bytes);
    if (int r = lane_barrier(ctx)) return r;
    HIPCHK(ctx, hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int icf_d2h(icf_ctx *ctx, void *dst, const void *src, size_t src_off, size_t bytes)
{
    if (!ctx || (!dst && bytes) || (!src && bytes)) return ICF_E_INVALID;
    if (ctx->capturing) return fail(ctx, ICF_E_CAPTURE, "host copies cannot be captured");
    if (int r = check_ctx(ctx)) return r;
    if (!bytes) return 0;
    if (int r = lane_barrier(ctx)) return r;
    HIPCHK(ctx, hipMemcpyAsync(dst, (const char *)src + src_off, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int icf_memset(icf_ctx *ctx, void *dst, int value, size_t bytes)
{
    if (!ctx || (!dst && bytes)) return ICF_E_INVALID;
    if (ctx->capturing) return fail(ctx, ICF_E_CAPTURE, "icf_memset during capture (clear buffers with a kernel)");
    if (int r = check_ctx(ctx)) return r;
    if (!bytes) return 0;
    invalidate_range(ctx, (uintptr_t)dst, (uintptr_t)dst + bytes);
    if (int r = lane_barrier(ctx)) return r;
    HIPCHK(ctx, hipMemsetAsync(dst, value, bytes, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int icf_synchronize(icf_ctx *ctx)
{
    if (int r = check_ctx(ctx)) return r;
    if (ctx->capturing) return fail(ctx, ICF_E_CAPTURE, "icf_synchronize during capture");
    if (int r = sync_lanes(ctx)) return r;
    prof_flush(ctx);
    return 0;
}

int icf_pattern_changed(icf_ctx *ctx)
{
    if (int r = check_ctx(ctx)) return r;
    if (ctx->capturing) return fail(ctx, ICF_E_CAPTURE, "icf_pattern_changed during capture");
    for (SlotMap &m : ctx->maps) m.suspect = true;
    ctx->adj.valid = false;
    ctx->patc.valid = false;
    ctx->epoch++;
    return 0;
}

double icf_last_device_ms(const icf_ctx *ctx) { return ctx ? ctx->last_ms : -1; }

int icf_d2d(icf_ctx *ctx, void *dst, size_t dst_off, const void *src, size_t src_off, size_t bytes)
{
    if (!ctx || (!dst && bytes) || (!src && bytes)) return ICF_E_INVALID;
    if (int r = check_ctx(ctx)) return r;
    if (!bytes) return 0;
    char *d = (char *)dst + dst_off;
    if (int r = note_write(ctx, d, bytes, "icf_d2d")) return r;
    OnLane ln(ctx, 0, [&] { return Access().rd((const char *)src + src_off, bytes).wr(d, bytes); });
    if (ln.err) return ln.err;
    Timer tm(ctx);
    LAUNCH(ctx, "k_copy", k_copy, dim3(blocks_for((long long)((bytes + 15) / 16), 256)), dim3(256), d,
           (const char *)src + src_off, bytes);
    if (int r = ln.done()) return r;
    return tm.finish();
}

int icf_probe_valu(icf_ctx *ctx, int fp64, int workgroups, int iters, void *out)
{
    if (!ctx || !out || workgroups <= 0 || iters <= 0) return ICF_E_INVALID;
    if (int r = check_ctx(ctx)) return r;
    const size_t bytes = (size_t)workgroups * 256 * (fp64 ? 8 : 4);
    if (int r = note_write(ctx, out, bytes, "icf_probe_valu")) return r;
    OnLane ln(ctx, 0, [&] { return Access().wr((const char *)out, bytes); });
    if (ln.err) return ln.err;
    Timer tm(ctx);
    // m slightly below one and c small keep every chain finite for any iteration count
    if (fp64)
        LAUNCH(ctx, "k_valu_probe<double>", k_valu_probe<double>, dim3(workgroups), dim3(256), (double *)out, iters,
               0.999999, 1e-6);
    else
        LAUNCH(ctx, "k_valu_probe<float>", k_valu_probe<float>, dim3(workgroups), dim3(256), (float *)out, iters,
               0.9999f, 1e-4f);
    if (int r = ln.done()) return r;
    return tm.finish();
}

int icf_set_blocking(icf_ctx *ctx, int blocking)
{
    if (!ctx) return ICF_E_INVALID;
    ctx->blocking = blocking != 0;
    return 0;
}

int icf_cz_counts(icf_ctx *ctx, int *nCZDamaged, int *nCZFailed)
{
    if (int r = check_ctx(ctx)) return r;
    if (ctx->capturing) return fail(ctx, ICF_E_CAPTURE, "icf_cz_counts during capture");
    if (int r = lane_barrier(ctx)) return r;
    int sums[2] = {0, 0};
    HIPCHK(ctx, hipMemcpyAsync(sums, ctx->small + 2, 2 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (nCZDamaged) *nCZDamaged = sums[0];
    if (nCZFailed) *nCZFailed = sums[1];
    return 0;
}

struct icf_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int device = 0;
    int ctx_id = 0;                  // the context that captured it ...
    unsigned long long epoch = 0;    // ... and its epoch at icf_graph_end
};

int icf_graph_begin(icf_ctx *ctx)
{
    if (int r = check_ctx(ctx)) return r;
    if (ctx->capturing) return fail(ctx, ICF_E_CAPTURE, "capture already in progress");
    if (int r = sync_lanes(ctx)) return r;
    prof_flush(ctx);
    HIPCHK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    ctx->capturing = true;
    return 0;
}

int icf_graph_end(icf_ctx *ctx, icf_graph **out)
{
    if (!ctx || !out) return ICF_E_INVALID;
    if (!ctx->capturing) return fail(ctx, ICF_E_CAPTURE, "no capture in progress");
    *out = nullptr;
    ctx->capturing = false;
    hipGraph_t g = nullptr;
    HIPCHK(ctx, hipStreamEndCapture(ctx->stream, &g));
    hipGraphExec_t ex = nullptr;
    hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(ctx, (int)e, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    icf_graph *gr = new icf_graph();
    gr->graph = g;
    gr->exec = ex;
    gr->device = ctx->device;
    gr->ctx_id = ctx->id;
    gr->epoch = ctx->epoch;
    *out = gr;
    return 0;
}

int icf_graph_launch(icf_ctx *ctx, icf_graph *g)
{
    if (int r = check_ctx(ctx)) return r;
    if (!g || g->device != ctx->device) return fail(ctx, ICF_E_INVALID, "icf_graph_launch: bad graph");
    if (ctx->capturing) return fail(ctx, ICF_E_CAPTURE, "icf_graph_launch during capture");
    if (g->ctx_id != ctx->id)
        return fail(ctx, ICF_E_CAPTURE, "icf_graph_launch: the graph was captured on another context");
    if (g->epoch != ctx->epoch)
        return fail(ctx, ICF_E_CAPTURE,
                    "icf_graph_launch: the graph is stale (since its capture, memory it uses was freed or moved, or "
                    "slot fields it was built from were written): capture it again");
    ConstUse cu(ctx);
    if (cu.err) return cu.err;
    OnLane ln(ctx, 0, everything);
    if (ln.err) return ln.err;
    Timer tm(ctx);
    HIPCHK(ctx, hipGraphLaunch(g->exec, ctx->stream));
    if (int r = ln.done()) return r;
    return tm.finish();
}

int icf_graph_destroy(icf_graph *g)
{
    if (!g) return ICF_E_INVALID;
    (void)hipSetDevice(g->device);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return 0;
}

int icf_set_profiling(icf_ctx *ctx, int on)
{
    if (!ctx) return ICF_E_INVALID;
    if (ctx->capturing) return fail(ctx, ICF_E_CAPTURE, "icf_set_profiling during capture");
    ctx->profiling = on != 0;
    return 0;
}

int icf_set_overlap(icf_ctx *ctx, int on)
{
    if (!ctx) return ICF_E_INVALID;
    if (ctx->capturing) return fail(ctx, ICF_E_CAPTURE, "icf_set_overlap during capture");
    ctx->overlap = on != 0;
    return 0;
}

int icf_set_autotune(icf_ctx *ctx, int rounds)
{
    if (!ctx || rounds < 0 || rounds > 1000) return ICF_E_INVALID;
    if (ctx->capturing) return fail(ctx, ICF_E_CAPTURE, "icf_set_autotune during capture");
    if (int r = check_ctx(ctx)) return r;
    std::vector<std::pair<icf_ctx::Tune *, int>> all;
    for (int a = 0; a < 2; a++)
        for (int f = 0; f < 2; f++) all.push_back({&ctx->tune[a][f], 2 * a + f});
    for (int r = 0; r < NSHAPE; r++) all.push_back({&ctx->shape[r], 4 + r});
    for (auto &tr : all) {
        icf_ctx::Tune &t = *tr.first;
        if (rounds == 0) {             // stop: fold what was timed, keep the fastest so far
            if (t.active) {
                tune_fold(ctx, t, tr.second);
                for (int c = 0; c < t.ncand; c++)
                    if (t.best[c] > 0 && (t.best[t.sel] == 0 || t.best[c] < t.best[t.sel])) t.sel = c;
            }
            t.active = false;
            continue;
        }
        for (hipEvent_t &e : t.ev)
            if (!e) HIPCHK(ctx, hipEventCreate(&e));
        t.active = true;
        t.fixed = false;
        t.turn = t.samples = 0;
        t.cand = -1;
        t.bucket = -1;
        for (float &b : t.best) b = 0;
    }
    ctx->tune_rounds = rounds;
    return 0;
}

int icf_get_autotune(icf_ctx *ctx, int *keys, int *stage, int *tuning, float *best_ms)
{
    if (!ctx) return ICF_E_INVALID;
    for (int r = 0; r < 4; r++) {
        icf_ctx::Tune &t = ctx->tune[r >> 1][r & 1];
        if (t.active && !ctx->capturing) tune_fold(ctx, t, r);   // a finished last timing counts
        if (keys) keys[r] = cand_keys(t.sel);
        if (stage) stage[r] = cand_stage(t.sel);
        if (tuning) tuning[r] = t.active ? 1 : 0;
        if (best_ms)
            for (int c = 0; c < NCAND; c++) best_ms[r * NCAND + c] = t.best[c];
    }
    return 0;
}

int icf_set_chunk_keys(icf_ctx *ctx, int accumulate, int fp64, int keys, int stage)
{
    if (!ctx) return ICF_E_INVALID;
    icf_ctx::Tune &t = ctx->tune[accumulate != 0][fp64 != 0];
    if (stage == 0) stage = cand_stage(t.sel);
    int c = 0;
    while (c < NCAND && (cand_keys(c) != keys || cand_stage(c) != stage)) c++;
    if (c == NCAND)
        return fail(ctx, ICF_E_INVALID, "icf_set_chunk_keys: keys must be 64, 128, 256 or 512 and stage 2 or 3");
    if (t.active && !ctx->capturing) tune_fold(ctx, t, 2 * (accumulate != 0) + (fp64 != 0));
    t.active = false;
    t.fixed = true;
    t.cand = -1;
    t.sel = c;
    return 0;
}

int icf_set_shape(icf_ctx *ctx, int role, int value)
{
    if (!ctx) return ICF_E_INVALID;
    if (role < 0 || role >= NSHAPE) return fail(ctx, ICF_E_INVALID, "icf_set_shape: unknown role");
    int c = 0;
    while (c < shape_ncand(role) && shape_value(role, c) != value) c++;
    if (c == shape_ncand(role))
        return fail(ctx, ICF_E_INVALID, role == ICF_SHAPE_CLOSEST ? "icf_set_shape: lanes per tuple must be 4, 8 or 16"
                                        : role == ICF_SHAPE_ORDER ? "icf_set_shape: the order must be 0 or 1"
                                                                  : "icf_set_shape: workgroup size must be 64, 128 or 256");
    icf_ctx::Tune &t = ctx->shape[role];
    if (t.active && !ctx->capturing) tune_fold(ctx, t, 4 + role);
    t.active = false;
    t.fixed = true;
    t.cand = -1;
    t.sel = c;
    return 0;
}

int icf_get_shape(icf_ctx *ctx, int role, int *value, int *tuning, float *best_ms)
{
    if (!ctx) return ICF_E_INVALID;
    if (role < 0 || role >= NSHAPE) return fail(ctx, ICF_E_INVALID, "icf_get_shape: unknown role");
    icf_ctx::Tune &t = ctx->shape[role];
    if (t.active && !ctx->capturing) tune_fold(ctx, t, 4 + role);
    if (value) *value = shape_value(role, t.sel);
    if (tuning) *tuning = t.active ? 1 : 0;
    if (best_ms)
        for (int c = 0; c < NSHAPE_CAND; c++) best_ms[c] = c < t.ncand ? t.best[c] : 0.f;
    return 0;
}

int icf_profile_reset(icf_ctx *ctx)
{
    if (!ctx) return ICF_E_INVALID;
    ctx->prof_names.clear();
    ctx->prof_ms.clear();
    ctx->prof_n.clear();
    return 0;
}

int icf_profile_count(const icf_ctx *ctx) { return ctx ? (int)ctx->prof_names.size() : ICF_E_INVALID; }

int icf_profile_entry(const icf_ctx *ctx, int i, char *name, int cap, double *total_ms, long long *launches)
{
    if (!ctx || i < 0 || i >= (int)ctx->prof_names.size()) return ICF_E_INVALID;
    if (name && cap > 0) {
        strncpy(name, ctx->prof_names[i].c_str(), (size_t)cap - 1);
        name[cap - 1] = 0;
    }
    if (total_ms) *total_ms = ctx->prof_ms[i];
    if (launches) *launches = ctx->prof_n[i];
    return 0;
}

int icf_elem_force(icf_ctx *ctx, const int *ie, double *dn, double h, double *A, double *b, int nElems, int el_stride,
                   int nd_stride)
{
    if (int r = check_ctx(ctx)) return r;
    if (bad_units(nElems, el_stride) || nd_stride < 0 || (nElems && (!ie || !dn || !A || !b)) || !(h > 0))
        return fail(ctx, ICF_E_INVALID, "icf_elem_force: invalid argument");
    ConstUse cu(ctx);
    if (cu.err) return cu.err;
    if (nElems == 0) return 0;
    OnLane ln(ctx, 0, everything);
    if (ln.err) return ln.err;
    SlotMap *mv, *mr, *mf;
    std::vector<SrcKey> sv{{KIND_TET, ie, nElems, el_stride, EL_SLOTS, 16}};
    std::vector<SrcKey> sr{{KIND_TET, ie, nElems, el_stride, EL_ROWS, 4}};
    std::vector<SrcKey> sfv{{KIND_TET, ie, nElems, el_stride, EL_NODES, 4}};
    if (int r = get_map(ctx, TGT_VALS, -1, sv, &mv)) return r;
    if (int r = get_map(ctx, TGT_RHS, -1, sr, &mr)) return r;
    if (int r = get_map(ctx, TGT_FORCE, -1, sfv, &mf)) return r;
    Timer tm(ctx);
    if (int r = stage_tets(ctx, ie, dn, nElems, el_stride, nd_stride)) return r;
    Stage st = make_stage(ctx, h);
    launch_vec3(ctx, mf, dn + (size_t)nd_stride * ND_F, nd_stride, 1, 1, st);
    launch_vals(ctx, mv, A, 1, st);
    launch_vec3(ctx, mr, b, 1, 3, 1, st);
    if (int r = ln.done()) return r;
    return tm.finish();
}

int icf_cz_force(icf_ctx *ctx, double *dcz, int *icz, double *dn, double *A, double *b, double h, int nCZs,
                 int cz_stride, int nd_stride)
{
    if (int r = check_ctx(ctx)) return r;
    if (bad_units(nCZs, cz_stride) || nd_stride < 0 || (nCZs && (!dcz || !icz || !dn || !A || !b)))
        return fail(ctx, ICF_E_INVALID, "icf_cz_force: invalid argument");
    ConstUse cu(ctx);
    if (cu.err) return cu.err;
    if (nCZs == 0) return 0;
    OnLane ln(ctx, 0, everything);
    if (ln.err) return ln.err;
    SlotMap *mv, *mr;
    std::vector<SrcKey> sv{{KIND_CZ, icz, nCZs, cz_stride, CZI_SLOTS, 36}};
    std::vector<SrcKey> sr{{KIND_CZ, icz, nCZs, cz_stride, CZI_ROWS, 6}};
    if (int r = get_map(ctx, TGT_VALS, -1, sv, &mv)) return r;
    if (int r = get_map(ctx, TGT_RHS, -1, sr, &mr)) return r;
    Timer tm(ctx);
    if (int r = stage_czs(ctx, dcz, icz, dn, nCZs, cz_stride, nd_stride)) return r;
    Stage st = make_stage(ctx, h);
    launch_vals(ctx, mv, A, 1, st);
    launch_vec3(ctx, mr, b, 1, 3, 1, st);
    if (int r = ln.done()) return r;
    return tm.finish();
}

int icf_narrow_phase(icf_ctx *ctx, int nPairs, const double *dn, const int *ie, int *narrowList, int el_all_stride,
                     int nd_stride)
{
    if (int r = check_ctx(ctx)) return r;
    if (nPairs < 0 || el_all_stride < 0 || nd_stride < 0 || (nPairs && (!dn || !ie || !narrowList)))
        return fail(ctx, ICF_E_INVALID, "icf_narrow_phase: invalid argument");
    if (nPairs == 0) return 0;
    if (int r = note_write(ctx, narrowList, sizeof(int) * 2 * (size_t)nPairs, "icf_narrow_phase")) return r;
    // reads node positions and the tets' node ids, rewrites the list
    OnLane ln(ctx, contact_lane(ctx), [&] {
        return Access()
            .rd(dn + (size_t)nd_stride * ND_XC, 3 * (size_t)nd_stride)
            .rd(ie, 4 * (size_t)el_all_stride)
            .wr(narrowList, 2 * (size_t)nPairs);
    });
    if (ln.err) return ln.err;
    Timer tm(ctx);
    const int B = shape_begin(ctx, ICF_SHAPE_NARROW, nPairs);
    LAUNCH(ctx, "k_narrow_phase", k_narrow_phase, dim3(blocks_for(nPairs, B)), dim3(B), nPairs, dn, ie,
           narrowList, el_all_stride, nd_stride);
    shape_end(ctx, ICF_SHAPE_NARROW);
    if (int r = ln.done()) return r;
    return tm.finish();
}

// The compact adjacency lists of ie (valid on return), built synchronously when missing.
int get_adjacency(icf_ctx *ctx, const int *ie, int es)
{
    icf_ctx::Adj &a = ctx->adj;
    if (a.valid && a.ie == ie && a.es == es) return 0;
    if (ctx->capturing)
        return fail(ctx, ICF_E_CAPTURE, "the adjacency lists must be built: run the call once uncaptured");
    void *base = nullptr;
    size_t sz = 0;
    HIPCHK(ctx, hipMemGetAddressRange(&base, &sz, const_cast<int *>(ie)));
    const uintptr_t end = (uintptr_t)base + sz;
    const long long rows = ((long long)(end - (uintptr_t)ie)) / (4LL * std::max(es, 1));
    if (rows < 5) return fail(ctx, ICF_E_INVALID, "icf_find_closest_face: ie holds fewer than 5 rows");
    const int max_rows = (int)std::min(rows - 5, 1LL << 20);
    if (int r = lane_barrier(ctx)) return r;
    if (int r = grow(ctx, a.off, sizeof(int) * ((size_t)es + 1))) return r;
    int *off = (int *)a.off.p;
    size_t tb = 0;
    HIPCHK(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, tb, off, off, es + 1, ctx->stream));
    if (int r = grow(ctx, ctx->scan_tmp, tb)) return r;
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, ctx->stream, off + es, 1, 0);
    hipLaunchKernelGGL(k_adj_count, dim3(blocks_for(es, 256)), dim3(256), 0, ctx->stream, ie, es, max_rows, off);
    HIPCHK(ctx, hipcub::DeviceScan::ExclusiveSum(ctx->scan_tmp.p, tb, off, off, es + 1, ctx->stream));
    int total = 0;
    HIPCHK(ctx, hipMemcpyAsync(&total, off + es, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (int r = grow(ctx, a.ids, sizeof(int) * (size_t)std::max(total, 1))) return r;
    hipLaunchKernelGGL(k_adj_fill, dim3(blocks_for(es, 256)), dim3(256), 0, ctx->stream, ie, es, off, (int *)a.ids.p);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    a.ie = ie;
    a.es = es;
    a.lo = (uintptr_t)base;
    a.hi = end;
    a.valid = true;
    return 0;
}

// k_closest_face with the tuned lanes per tuple, on `stream` (the profiled launch when it is
// the context's current stream)
void launch_closest_face(icf_ctx *ctx, int nPairs, int ts, int *nl2, const double *dn, const int *adj_off,
                         const int *adj_ids, const int *faces, int ns, int fs, hipStream_t stream)
{
    const bool timed = stream == ctx->cur;   // tuning times launches on the current stream only
    const int G = shape_begin(ctx, ICF_SHAPE_CLOSEST, nPairs, !timed);
    const dim3 g(blocks_for((long long)nPairs * G, 256)), b(256);
    auto go = [&](auto kern) {
        if (stream == ctx->cur) LAUNCH(ctx, "k_closest_face", kern, g, b, nPairs, ts, nl2, dn, adj_off, adj_ids, faces, ns, fs);
        else hipLaunchKernelGGL(kern, g, b, 0, stream, nPairs, ts, nl2, dn, adj_off, adj_ids, faces, ns, fs);
    };
    if (G == 4) go(k_closest_face<4>);
    else if (G == 16) go(k_closest_face<16>);
    else go(k_closest_face<8>);
    if (timed) shape_end(ctx, ICF_SHAPE_CLOSEST);
}

int icf_find_closest_face(icf_ctx *ctx, int nPairs, int tet_stride, int *nl2, const double *dn, const int *ie,
                          const int *faces, int nd_stride, int el_all_stride, int fc_stride)
{
    if (int r = check_ctx(ctx)) return r;
    if (bad_units(nPairs, tet_stride) || nd_stride < 0 || el_all_stride < 0 || fc_stride < 0 ||
        (nPairs && (!nl2 || !dn || !ie || !faces)))
        return fail(ctx, ICF_E_INVALID, "icf_find_closest_face: invalid argument");
    if (nPairs == 0) return 0;
    if (int r = note_write(ctx, nl2 + tet_stride, sizeof(int) * 3 * (size_t)tet_stride, "icf_find_closest_face"))
        return r;
    // the surface tets' adjacency lists, read from their compact copy
    if (int r = get_adjacency(ctx, ie, el_all_stride)) return r;
    OnLane ln(ctx, contact_lane(ctx), [&] {
        return Access()
            .wr(nl2, 4 * (size_t)tet_stride)
            .rd(dn + (size_t)nd_stride * ND_XC, 3 * (size_t)nd_stride)
            .rd((const char *)ctx->adj.off.p, ctx->adj.off.bytes)
            .rd((const char *)ctx->adj.ids.p, ctx->adj.ids.bytes)
            .rd(faces, 3 * (size_t)fc_stride);
    });
    if (ln.err) return ln.err;
    Timer tm(ctx);
    launch_closest_face(ctx, nPairs, tet_stride, nl2, dn, (const int *)ctx->adj.off.p, (const int *)ctx->adj.ids.p,
                        faces, nd_stride, fc_stride, ctx->cur);
    if (int r = ln.done()) return r;
    return tm.finish();
}

int icf_find_closest_face_noalloc(icf_ctx *ctx, int nPairs, int tet_stride, int *nl2, const double *dn,
                                  const int *ie, const int *faces, int nd_stride, int el_all_stride, int fc_stride)
{
    if (int r = check_ctx(ctx)) return r;
    if (bad_units(nPairs, tet_stride) || nd_stride < 0 || el_all_stride < 0 || fc_stride < 0 ||
        (nPairs && (!nl2 || !dn || !ie || !faces)))
        return fail(ctx, ICF_E_INVALID, "icf_find_closest_face_noalloc: invalid argument");
    if (nPairs == 0) return 0;
    if (!ctx->noalloc) return fail(ctx, ICF_E_INVALID, "icf_find_closest_face_noalloc: no scratch on this device");
    if (int r = note_write(ctx, nl2 + tet_stride, sizeof(int) * 3 * (size_t)tet_stride,
                           "icf_find_closest_face_noalloc"))
        return r;
    const int es = el_all_stride;
    // adjacent-face rows of ie, from the extent of its allocation (as icf_find_closest_face)
    void *base = nullptr;
    size_t sz = 0;
    HIPCHK(ctx, hipMemGetAddressRange(&base, &sz, const_cast<int *>(ie)));
    const long long rows = ((long long)((uintptr_t)base + sz - (uintptr_t)ie)) / (4LL * std::max(es, 1));
    if (rows < 5) return fail(ctx, ICF_E_INVALID, "icf_find_closest_face_noalloc: ie holds fewer than 5 rows");
    const int max_rows = (int)std::min(rows - 5, 1LL << 20);
    // scratch layout: list offsets [es + 1], the scan's temporary storage, the lists (at most
    // max_rows faces per tet: k_adj_count clamps the counts)
    auto up = [](size_t x) { return (x + 255) / 256 * 256; };
    size_t tb = 0;
    HIPCHK(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, tb, (int *)nullptr, (int *)nullptr, es + 1, ctx->stream));
    const size_t tmp_at = up(sizeof(int) * ((size_t)es + 1)), ids_at = tmp_at + up(tb);
    const size_t need = ids_at + sizeof(int) * (size_t)es * (size_t)std::max(max_rows, 1);
    if (need > NOALLOC_SCRATCH_BYTES)
        return fail(ctx, ICF_E_CAPACITY, "icf_find_closest_face_noalloc: el_all_stride " + std::to_string(es) +
                                             " x " + std::to_string(max_rows) + " adjacent-face rows need " +
                                             std::to_string(need) + " bytes of scratch, more than its " +
                                             std::to_string(NOALLOC_SCRATCH_BYTES) + "; use icf_find_closest_face");
    int *off = (int *)ctx->noalloc;
    void *tmp = ctx->noalloc + tmp_at;
    int *ids = (int *)(ctx->noalloc + ids_at);
    // after every earlier call of the context, on its stream; later calls of either lane wait
    if (int r = lane_barrier(ctx)) return r;
    Timer tm(ctx);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, ctx->stream, off + es, 1, 0);
    hipLaunchKernelGGL(k_adj_count, dim3(blocks_for(es, 256)), dim3(256), 0, ctx->stream, ie, es, max_rows, off);
    HIPCHK(ctx, hipcub::DeviceScan::ExclusiveSum(tmp, tb, off, off, es + 1, ctx->stream));
    hipLaunchKernelGGL(k_adj_fill, dim3(blocks_for(es, 256)), dim3(256), 0, ctx->stream, ie, es, (const int *)off,
                       ids);
    launch_closest_face(ctx, nPairs, tet_stride, nl2, dn, (const int *)off, (const int *)ids, faces, nd_stride,
                        fc_stride, ctx->stream);
    if (int r = lane_exit(ctx, 0, everything)) return r;
    return tm.finish();
}

int icf_collision_response(icf_ctx *ctx, const int *icr, double *dn, double h, double *A, double *b, int nImpacts,
                           int cr_stride, int nd_stride, double k, double distanceEpsilon)
{
    (void)h;
    if (int r = check_ctx(ctx)) return r;
    if (bad_units(nImpacts, cr_stride) || nd_stride < 0 || (nImpacts && (!icr || !dn || !A || !b)))
        return fail(ctx, ICF_E_INVALID, "icf_collision_response: invalid argument");
    if (nImpacts == 0) return 0;
    SlotMap *mv, *mr, *mf;
    std::vector<SrcKey> sv{{KIND_CONTACT, icr, nImpacts, cr_stride, CR_SLOTS, 16}};
    std::vector<SrcKey> sr{{KIND_CONTACT, icr, nImpacts, cr_stride, CR_ROWS, 4}};
    std::vector<SrcKey> sfv{{KIND_CONTACT, icr, nImpacts, cr_stride, CR_NODES, 4}};
    if (int r = get_map(ctx, TGT_VALS, -1, sv, &mv)) return r;
    if (int r = get_map(ctx, TGT_RHS, -1, sr, &mr)) return r;
    if (int r = get_map(ctx, TGT_FORCE, -1, sfv, &mf)) return r;
    const bool mixed = ctx->precision == ICF_MIXED;
    if (int r = grow(ctx, ctx->ct, (mixed ? sizeof(float) : sizeof(double)) * CT_TAN_REC * (size_t)nImpacts)) return r;
    if (int r = grow(ctx, ctx->ct_f, sizeof(double) * CT_F_REC * (size_t)nImpacts)) return r;
    Timer tm(ctx);
    const size_t tan_bytes = (mixed ? sizeof(float) : sizeof(double)) * CT_TAN_REC * (size_t)nImpacts;
    const size_t f_bytes = sizeof(double) * CT_F_REC * (size_t)nImpacts;
    // per-impact phase: node positions and the impact records in, impact blocks out
    OnLane l1(ctx, contact_lane(ctx), [&] {
        return Access()
            .rd(icr, (CR_ROWS + 4) * (size_t)cr_stride)
            .rd(dn + (size_t)nd_stride * ND_XC, 3 * (size_t)nd_stride)
            .wr((const char *)ctx->ct.p, tan_bytes)
            .wr((const char *)ctx->ct_f.p, f_bytes);
    });
    if (l1.err) return l1.err;
    const int B = shape_begin(ctx, ICF_SHAPE_IMPACT, nImpacts);
    const dim3 cg(blocks_for(nImpacts, B)), cb(B);
    if (mixed)
        LAUNCH(ctx, "k_contact_impact<mixed>", k_contact_impact<float>, cg, cb, icr, dn, nImpacts, cr_stride, nd_stride,
               k, distanceEpsilon, (float *)ctx->ct.p, (double *)ctx->ct_f.p);
    else
        LAUNCH(ctx, "k_contact_impact<fp64>", k_contact_impact<double>, cg, cb, icr, dn, nImpacts, cr_stride,
               nd_stride, k, distanceEpsilon, (double *)ctx->ct.p, (double *)ctx->ct_f.p);
    shape_end(ctx, ICF_SHAPE_IMPACT);
    if (int r = l1.done()) return r;
    // sums into node forces, BCSR blocks (at slots the call does not bound) and rhs rows
    OnLane l0(ctx, 0, [&] {
        return Access()
            .rd((const char *)ctx->ct.p, tan_bytes)
            .rd((const char *)ctx->ct_f.p, f_bytes)
            .rd(icr, (CR_ROWS + 4) * (size_t)cr_stride)
            .wr(dn + (size_t)nd_stride * ND_F, 3 * (size_t)nd_stride)
            .alloc(A, true)
            .alloc(b, true);
    });
    if (l0.err) return l0.err;
    Stage st = make_stage(ctx, 1.0);
    double *F = dn + (size_t)nd_stride * ND_F;
    if (!(ctx->overlap && contact_sums_fused(ctx, mf, F, nd_stride, mv, A, mr, b, st))) {
        launch_vec3(ctx, mf, F, nd_stride, 1, 1, st);
        launch_vals(ctx, mv, A, 1, st, true);
        launch_vec3(ctx, mr, b, 1, 3, 1, st);
    }
    if (int r = l0.done()) return r;
    return tm.finish();
}

int icf_sum_int(icf_ctx *ctx, const int *d, int n, int *out)
{
    if (int r = check_ctx(ctx)) return r;
    if (n < 0 || !out || (n && !d)) return fail(ctx, ICF_E_INVALID, "icf_sum_int: invalid argument");
    if (ctx->capturing) return fail(ctx, ICF_E_CAPTURE, "icf_sum_int returns a host value");
    *out = 0;
    if (n == 0) return 0;
    if (int r = lane_barrier(ctx)) return r;
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, ctx->stream, ctx->small + 4, 1, 0);
    hipLaunchKernelGGL(k_sum_int, dim3(n < 64 * 256 ? blocks_for(n, 256) : 64), dim3(256), 0, ctx->stream, d, n,
                       ctx->small + 4);
    HIPCHK(ctx, hipMemcpyAsync(out, ctx->small + 4, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

// The fused call; map_ids (optional) receives the ids of its value and rhs maps.
static int assemble_impl(icf_ctx *ctx, const int *ie, int nElems, int el_stride, double *dcz, int *icz, int nCZs,
                         int cz_stride, double *dn, int nd_stride, double h, double *A, int nnzb, double *b, int N,
                         int *nCZDamaged, int *nCZFailed, int *map_ids)
{
    if (int r = check_ctx(ctx)) return r;
    if (bad_units(nElems, el_stride) || bad_units(nCZs, cz_stride) || nd_stride < 0 || nnzb < 0 || N < 0 ||
        (nElems && !ie) || (nCZs && (!dcz || !icz)) || !dn || (nnzb && !A) || (N && !b) || !(h > 0))
        return fail(ctx, ICF_E_INVALID, "icf_assemble_elems_and_czs: invalid argument");
    ConstUse cu(ctx);
    if (cu.err) return cu.err;
    // node fields read (x0, u, v, a, x) and written (F); CZ records rewritten in place
    const size_t ns = (size_t)nd_stride;
    OnLane ln(ctx, 0, [&] {
        return Access()
            .rd(ie, (EL_ROWS + 4) * (size_t)el_stride)
            .alloc(nCZs ? dcz : nullptr, true)
            .alloc(nCZs ? icz : nullptr, true)
            .rd(dn, ND_F * ns)
            .wr(dn + ND_F * ns, 3 * ns)
            .rd(dn + ND_XC * ns, 3 * ns)
            .wr(A, 9 * (size_t)nnzb)
            .wr(b, 3 * (size_t)N);
    });
    if (ln.err) return ln.err;
    std::vector<SrcKey> sv, sr, sfv;
    if (nElems) {
        sv.push_back({KIND_TET, ie, nElems, el_stride, EL_SLOTS, 16});
        sr.push_back({KIND_TET, ie, nElems, el_stride, EL_ROWS, 4});
        sfv.push_back({KIND_TET, ie, nElems, el_stride, EL_NODES, 4});
    }
    if (nCZs) {
        sv.push_back({KIND_CZ, icz, nCZs, cz_stride, CZI_SLOTS, 36});
        sr.push_back({KIND_CZ, icz, nCZs, cz_stride, CZI_ROWS, 6});
    }
    SlotMap *mv, *mr, *mf = nullptr;
    if (int r = get_map(ctx, TGT_VALS, nnzb, sv, &mv)) return r;
    if (int r = get_map(ctx, TGT_RHS, N, sr, &mr)) return r;
    if (nElems)
        if (int r = get_map(ctx, TGT_FORCE, -1, sfv, &mf)) return r;
    if (map_ids) {
        map_ids[0] = mv->id;
        map_ids[1] = mr->id;
    }
    bool tsum = false;
    if (mf) {
        MapSrc tet{ie, nElems, el_stride, EL_NODES, 4, KIND_TET};
        if (int r = get_tsum_link(ctx, mr, mf, tet, &tsum)) return r;
        if (tsum)
            if (int r = grow(ctx, ctx->tsum, sizeof(double) * 3 * (size_t)mf->nkeys)) return r;
    }
    // buffers grow before any work is issued (growing synchronizes the context's stream)
    if (int r = stage_tets(ctx, ie, dn, 0, el_stride, nd_stride, nElems)) return r;
    if (int r = stage_czs(ctx, dcz, icz, dn, 0, cz_stride, nd_stride, nCZs)) return r;
    Timer tm(ctx);
    int *counts = ctx->small + 2;   // the two counts of GPU_Functionality.cs:474-475
    const int *dmg = icz ? icz + (size_t)cz_stride * CZI_TDAMAGED : nullptr;
    const int *fld = icz ? icz + (size_t)cz_stride * CZI_TFAILED : nullptr;
    const dim3 cnt_grid(nCZs < 64 * 256 ? blocks_for(std::max(nCZs, 1), 256) : 64);
    const bool fuse = ctx->overlap;
    // phase 1: tets and CZs (the counters are cleared by a kernel, so the call can be captured)
    if (fuse) {
        phase1_fused(ctx, ie, dn, nElems, el_stride, nd_stride, dcz, icz, nCZs, cz_stride, counts);
    } else {
        LAUNCH(ctx, "k_fill_int", k_fill_int, dim3(1), dim3(64), counts, 2, 0);
        if (int r = stage_tets(ctx, ie, dn, nElems, el_stride, nd_stride)) return r;
        if (int r = stage_czs(ctx, dcz, icz, dn, nCZs, cz_stride, nd_stride)) return r;
    }
    Stage st = make_stage(ctx, h);
    double *F = dn + (size_t)nd_stride * ND_F;
    if (tsum && fuse && vals_force_rhs_fused(ctx, mv, A, mf, F, nd_stride, mr, b, dmg, fld, nCZs, counts, st)) {
        // done: node forces with the rhs rows, the counts and the BCSR values in one launch
    } else if (tsum) {
        // BCSR values and node forces (keeping each node's tet sum), then the rhs rows from
        // those sums plus their CZ terms, and the counts
        if (!(fuse && vals_force_fused(ctx, mv, A, mf, F, nd_stride, st))) {
            launch_vals(ctx, mv, A, 0, st);
            LAUNCH(ctx, "k_reduce_vec3<force,acc>", (k_reduce_vec3<true, true>), dim3(blocks_for(mf->nlaunch, 256)),
                   dim3(256), mf->off, mf->codes, mf->keys, mf->nlaunch, F, nd_stride, 1, st, (double *)ctx->tsum.p);
        }
        if (fuse && (N || nCZs)) {
            const int nbr = N ? (int)(blocks_for(N, 256) + 7) / 8 * 8 : 0;
            LAUNCH(ctx, "k_rhs_counts", k_rhs_counts_fused, dim3(nbr + (nCZs ? cnt_grid.x : 0)), dim3(256), nbr,
                   mr->off, mr->codes, mr->cz_begin, mr->row_node, (const double *)ctx->tsum.p, N, b, st, dmg, fld,
                   nCZs, counts);
        } else {
            if (N)
                LAUNCH(ctx, "k_reduce_rhs_tsum", k_reduce_rhs_tsum, dim3(blocks_for(N, 256)), dim3(256), mr->off,
                       mr->codes, mr->cz_begin, mr->row_node, (const double *)ctx->tsum.p, N, b, st);
            if (nCZs) LAUNCH(ctx, "k_count_flags", k_count_flags, cnt_grid, dim3(256), dmg, fld, nCZs, counts);
        }
    } else {
        launch_vals(ctx, mv, A, 0, st);
        if (mf) launch_vec3(ctx, mf, F, nd_stride, 1, 1, st);
        launch_vec3(ctx, mr, b, 1, 3, 0, st);
        if (nCZs) LAUNCH(ctx, "k_count_flags", k_count_flags, cnt_grid, dim3(256), dmg, fld, nCZs, counts);
    }
    if (int r = ln.done()) return r;
    if (int r = tm.finish()) return r;
    if ((nCZDamaged || nCZFailed) && ctx->blocking && !ctx->capturing) {
        int sums[2] = {0, 0};
        HIPCHK(ctx, hipMemcpyAsync(sums, ctx->small + 2, 2 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        if (nCZDamaged) *nCZDamaged = sums[0];
        if (nCZFailed) *nCZFailed = sums[1];
    }
    return 0;
}

int icf_assemble_elems_and_czs(icf_ctx *ctx, const int *ie, int nElems, int el_stride, double *dcz, int *icz, int nCZs,
                               int cz_stride, double *dn, int nd_stride, double h, double *A, int nnzb, double *b,
                               int N, int *nCZDamaged, int *nCZFailed)
{
    return assemble_impl(ctx, ie, nElems, el_stride, dcz, icz, nCZs, cz_stride, dn, nd_stride, h, A, nnzb, b, N,
                         nCZDamaged, nCZFailed, nullptr);
}

int icf_assemble_batched(icf_ctx *ctx, const icf_worklist *wl, const int *ie, int el_stride, double *dcz, int *icz,
                         int cz_stride, double *dn, int nd_stride, double h, double *A, double *b, int *nCZDamaged,
                         int *nCZFailed)
{
    if (int r = check_ctx(ctx)) return r;
    if (!wl || wl->nSpecimens < 1 || wl->nSpecimens > (1 << 24) || !wl->tet_off || !wl->cz_off || !wl->row_off ||
        !wl->nnzb_off)
        return fail(ctx, ICF_E_INVALID, "icf_assemble_batched: invalid worklist");
    const int n = wl->nSpecimens;
    const int *tabs[4] = {wl->tet_off, wl->cz_off, wl->row_off, wl->nnzb_off};
    std::vector<int> off(4 * ((size_t)n + 1));
    for (int k = 0; k < 4; k++) {
        if (tabs[k][0] != 0) return fail(ctx, ICF_E_INVALID, "icf_assemble_batched: offsets must start at 0");
        for (int s = 0; s <= n; s++) {
            if (s && tabs[k][s] < tabs[k][s - 1])
                return fail(ctx, ICF_E_INVALID, "icf_assemble_batched: offsets must not decrease");
            off[(size_t)k * (n + 1) + s] = tabs[k][s];
        }
    }
    const int nE = wl->tet_off[n], nC = wl->cz_off[n], N = wl->row_off[n], nnzb = wl->nnzb_off[n];
    icf_ctx::Batch &B = ctx->batch;
    if (off != B.off) {
        if (ctx->capturing)
            return fail(ctx, ICF_E_CAPTURE, "icf_assemble_batched: new offsets cannot be uploaded during capture");
        if (int r = lane_barrier(ctx)) return r;
        if (int r = grow(ctx, B.d, sizeof(int) * (off.size() + 2 * (size_t)n))) return r;
        HIPCHK(ctx, hipMemcpyAsync(B.d.p, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        B.off = off;
        B.n = n;
        B.checked.clear();
        ctx->epoch++;              // a captured batch reads the old offsets
    }
    const int *d = (const int *)B.d.p;
    const BatchOff bo{d, d + (n + 1), d + 2 * (n + 1), d + 3 * (n + 1), n};
    int *cnt = (int *)B.d.p + 4 * (n + 1);
    // The block-diagonal layout, checked before anything is written whenever the units or the
    // slot fields may have changed (a library write to them moves the epoch).
    const std::vector<long long> sig{(long long)(uintptr_t)ie, (long long)(uintptr_t)icz, el_stride, cz_stride, nE,
                                     nC, (long long)ctx->epoch};
    if (sig != B.checked) {
        if (ctx->capturing)
            return fail(ctx, ICF_E_CAPTURE, "icf_assemble_batched: run the batch once uncaptured before capture");
        if (int r = lane_barrier(ctx)) return r;
        int *d_err = ctx->small + 1;
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, ctx->stream, d_err, 1, 0);
        if (nE + nC)
            hipLaunchKernelGGL(k_batch_check, dim3(blocks_for((long long)nE + nC, 256)), dim3(256), 0, ctx->stream, ie,
                               nE, el_stride, icz, nC, cz_stride, bo, d_err);
        int e = 0;
        HIPCHK(ctx, hipGetLastError());
        HIPCHK(ctx, hipMemcpyAsync(&e, d_err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        if (e)
            return fail(ctx, ICF_E_INDEX,
                        "icf_assemble_batched: a unit writes a BCSR block or rhs row outside its specimen's ranges");
    }
    B.cz_off = bo.cz;
    B.spec_counts = cnt;
    B.spec_counted = false;
    const int ra = assemble_impl(ctx, ie, nE, el_stride, dcz, icz, nC, cz_stride, dn, nd_stride, h, A, nnzb, b, N,
                                 nullptr, nullptr, nullptr);
    B.cz_off = nullptr;
    B.spec_counts = nullptr;
    if (ra) return ra;
    if (sig != B.checked) {
        B.checked = sig;
        B.checked.back() = (long long)ctx->epoch;   // map checks inside the call may have moved it
    }
    OnLane ln(ctx, 0, [&] {
        return Access()
            .rd(nE ? ie : nullptr, (EL_ROWS + 4) * (size_t)el_stride)
            .rd(nC ? icz : nullptr, (CZI_ROWS + 6) * (size_t)cz_stride);
    });
    if (ln.err) return ln.err;
    if (nC && B.spec_counted) {
        // written by the fused value launch
    } else if (nC) {
        const int *dmg = icz + (size_t)cz_stride * CZI_TDAMAGED, *fld = icz + (size_t)cz_stride * CZI_TFAILED;
        LAUNCH(ctx, "k_batch_counts", k_batch_counts, dim3(n), dim3(256), dmg, fld, bo.cz, cnt);
    } else {
        LAUNCH(ctx, "k_fill_int", k_fill_int, dim3(blocks_for(2 * (long long)n, 256)), dim3(256), cnt, 2 * n, 0);
    }
    if (int r = ln.done()) return r;
    if ((nCZDamaged || nCZFailed) && ctx->blocking && !ctx->capturing) {
        std::vector<int> c(2 * (size_t)n);
        HIPCHK(ctx, hipMemcpyAsync(c.data(), cnt, sizeof(int) * c.size(), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        for (int k = 0; k < n; k++) {
            if (nCZDamaged) nCZDamaged[k] = c[2 * (size_t)k];
            if (nCZFailed) nCZFailed[k] = c[2 * (size_t)k + 1];
        }
    }
    return 0;
}

int icf_batch_cz_counts(icf_ctx *ctx, int *nCZDamaged, int *nCZFailed)
{
    if (int r = check_ctx(ctx)) return r;
    if (ctx->capturing) return fail(ctx, ICF_E_CAPTURE, "icf_batch_cz_counts during capture");
    const int n = ctx->batch.n;
    if (n == 0) return fail(ctx, ICF_E_INVALID, "icf_batch_cz_counts: no batched call yet");
    if (int r = sync_lanes(ctx)) return r;
    std::vector<int> c(2 * (size_t)n);
    HIPCHK(ctx, hipMemcpyAsync(c.data(), (int *)ctx->batch.d.p + 4 * (n + 1), sizeof(int) * c.size(),
                               hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < n; k++) {
        if (nCZDamaged) nCZDamaged[k] = c[2 * (size_t)k];
        if (nCZFailed) nCZFailed[k] = c[2 * (size_t)k + 1];
    }
    return 0;
}

int icf_nodes_infer(icf_ctx *ctx, const double *st, int st_stride, double *dn, int nd_stride, int nNodes, double h,
                    double beta, double gamma)
{
    if (int r = check_ctx(ctx)) return r;
    if (bad_units(nNodes, st_stride) || bad_units(nNodes, nd_stride) || (nNodes && (!st || !dn)) || !(h > 0) ||
        !(beta > 0))
        return fail(ctx, ICF_E_INVALID, "icf_nodes_infer: invalid argument");
    if (nNodes == 0) return 0;
    OnLane ln(ctx, 0, everything);
    if (ln.err) return ln.err;
    Timer tm(ctx);
    LAUNCH(ctx, "k_nodes_infer", k_nodes_infer, dim3(blocks_for(3LL * nNodes, 256)), dim3(256), st, st_stride, dn,
           nd_stride, nNodes, h, beta, gamma);
    if (int r = ln.done()) return r;
    return tm.finish();
}

int icf_nodes_add_dx(icf_ctx *ctx, double *st, int st_stride, const int *altId, int nNodes, const double *dx)
{
    if (int r = check_ctx(ctx)) return r;
    if (bad_units(nNodes, st_stride) || (nNodes && (!st || !altId || !dx)))
        return fail(ctx, ICF_E_INVALID, "icf_nodes_add_dx: invalid argument");
    if (nNodes == 0) return 0;
    OnLane ln(ctx, 0, everything);
    if (ln.err) return ln.err;
    Timer tm(ctx);
    LAUNCH(ctx, "k_nodes_add_dx", k_nodes_add_dx, dim3(blocks_for(3LL * nNodes, 256)), dim3(256), st, st_stride,
           altId, nNodes, dx);
    if (int r = ln.done()) return r;
    return tm.finish();
}

int icf_nodes_accept(icf_ctx *ctx, double *st, int st_stride, const double *dn, int nd_stride, int nNodes, double h)
{
    if (int r = check_ctx(ctx)) return r;
    if (bad_units(nNodes, st_stride) || bad_units(nNodes, nd_stride) || (nNodes && (!st || !dn)) || !(h > 0))
        return fail(ctx, ICF_E_INVALID, "icf_nodes_accept: invalid argument");
    if (nNodes == 0) return 0;
    OnLane ln(ctx, 0, everything);
    if (ln.err) return ln.err;
    Timer tm(ctx);
    LAUNCH(ctx, "k_nodes_accept", k_nodes_accept, dim3(blocks_for(3LL * nNodes, 256)), dim3(256), st, st_stride, dn,
           nd_stride, nNodes, h);
    if (int r = ln.done()) return r;
    return tm.finish();
}

int icf_build_pattern(icf_ctx *ctx, const int *altId, int nNodes, const int *ie, int nElems, int el_stride,
                      const int *icz, int nCZs, int cz_stride, const int *icr, int nImpacts, int cr_stride, int N,
                      int *csr_rows, int *csr_cols, int max_nnz, int *nnz)
{
    if (int r = check_ctx(ctx)) return r;
    if (!ctx->capturing)
        if (int r = lane_barrier(ctx)) return r;
    if (!nnz || nNodes < 0 || N < 0 || max_nnz < 0 || bad_units(nElems, el_stride) || bad_units(nCZs, cz_stride) ||
        bad_units(nImpacts, cr_stride) || (nNodes && !altId) || (nElems && !ie) || (nCZs && !icz) ||
        (nImpacts && !icr) || !csr_rows || (max_nnz && !csr_cols))
        return fail(ctx, ICF_E_INVALID, "icf_build_pattern: invalid argument");
    if (ctx->capturing) return fail(ctx, ICF_E_CAPTURE, "icf_build_pattern returns a host value");
    *nnz = 0;
    // the static part (elements, CZs) and the impacts; the static part's distinct keys are kept
    // and reused while its node ids and the altIds are unchanged, so that a call with only new
    // impacts sorts the impact keys and merges them in
    PatSrc src[3] = {{ie, nElems, el_stride, EL_NODES, 4, 0, 0},
                     {icz, nCZs, cz_stride, CZI_VRTS, 6, 0, (long long)nElems * pat_pairs(4, 0)},
                     {icr, nImpacts, cr_stride, CR_NODES, 4, 1, 0}};
    const long long Ms = (long long)nElems * pat_pairs(4, 0) + (long long)nCZs * pat_pairs(6, 0);
    const long long Mi = (long long)nImpacts * pat_pairs(4, 1);
    if (Ms + Mi >= (1LL << 31) - 1)
        return fail(ctx, ICF_E_CAPACITY, "icf_build_pattern: 2^31 or more candidate pairs");
    int *d_err = ctx->small + 1, *d_diff = ctx->small + 2;
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, ctx->stream, d_err, 2, 0);
    int cb = 1;
    while ((1LL << cb) <= (long long)N) cb++;
    const PatFields pf{altId, nNodes, ie, nElems, el_stride, icz, nCZs, cz_stride};
    const long long nsnap = (long long)nNodes + 4LL * nElems + 6LL * nCZs;
    auto &pc = ctx->patc;
    bool hit = pc.valid && pc.nNodes == nNodes && pc.nElems == nElems && pc.el_stride == el_stride &&
               pc.nCZs == nCZs && pc.cz_stride == cz_stride && pc.N == N && pc.cb == cb;
    if (hit && nsnap > 0) {
        hipLaunchKernelGGL(k_pat_snap, dim3(blocks_for(nsnap, 256)), dim3(256), 0, ctx->stream, pf, nsnap,
                           (int *)pc.snap.p, 1, d_diff);
        int diff = 0;
        HIPCHK(ctx, hipGetLastError());
        HIPCHK(ctx, hipMemcpyAsync(&diff, d_diff, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        hit = diff == 0;
    }
    // sorted keys -> their distinct blocks, in order (flag, scan, compact); returns the count
    auto distinct = [&](const unsigned long long *keys, long long M, unsigned long long *out, int *count) -> int {
        if (int r = grow(ctx, ctx->bp_l, sizeof(int) * 2 * ((size_t)M + 1))) return r;
        int *flag = (int *)ctx->bp_l.p, *pos = flag + M + 1;
        size_t ub = 0;
        HIPCHK(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, ub, flag, pos, (int)M + 1, ctx->stream));
        if (int r = grow(ctx, ctx->scan_tmp, ub)) return r;
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, ctx->stream, flag + M, 1, 0);
        hipLaunchKernelGGL(k_pat_flag, dim3(blocks_for(M, 256)), dim3(256), 0, ctx->stream, keys, (int)M, cb, flag);
        HIPCHK(ctx, hipcub::DeviceScan::ExclusiveSum(ctx->scan_tmp.p, ub, flag, pos, (int)M + 1, ctx->stream));
        hipLaunchKernelGGL(k_pat_compact, dim3(blocks_for(M, 256)), dim3(256), 0, ctx->stream, keys, (int)M, flag,
                           pos, out);
        HIPCHK(ctx, hipGetLastError());
        HIPCHK(ctx, hipMemcpyAsync(count, pos + M, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        return 0;
    };
    auto sort_keys = [&](unsigned long long *in, unsigned long long *out, long long M) -> int {
        size_t sb = 0;
        HIPCHK(ctx, hipcub::DeviceRadixSort::SortKeys(nullptr, sb, in, out, (int)M, 0, 2 * cb, ctx->stream));
        if (int r = grow(ctx, ctx->scan_tmp, sb)) return r;
        HIPCHK(ctx, hipcub::DeviceRadixSort::SortKeys(ctx->scan_tmp.p, sb, in, out, (int)M, 0, 2 * cb, ctx->stream));
        return 0;
    };
    if (!hit) {
        pc.valid = false;
        int nk = 0;
        if (Ms > 0) {
            if (int r = grow(ctx, ctx->pat_a, sizeof(unsigned long long) * (size_t)Ms)) return r;
            if (int r = grow(ctx, ctx->pat_b, sizeof(unsigned long long) * (size_t)Ms)) return r;
            if (int r = grow(ctx, pc.keys, sizeof(unsigned long long) * (size_t)Ms)) return r;
            auto *a = (unsigned long long *)ctx->pat_a.p, *b = (unsigned long long *)ctx->pat_b.p;
            for (int k = 0; k < 2; k++)
                if (src[k].n)
                    hipLaunchKernelGGL(k_pat_emit, dim3(blocks_for((long long)src[k].n * pat_pairs(src[k].K, 0), 256)),
                                       dim3(256), 0, ctx->stream, src[k], altId, nNodes, N, cb, a, d_err);
            if (int r = sort_keys(a, b, Ms)) return r;
            if (int r = distinct(b, Ms, (unsigned long long *)pc.keys.p, &nk)) return r;
        }
        int err = 0;
        HIPCHK(ctx, hipMemcpyAsync(&err, d_err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        if (err) return fail(ctx, ICF_E_INDEX, "icf_build_pattern: a node id or altId is out of range");
        if (nsnap > 0) {
            if (int r = grow(ctx, pc.snap, sizeof(int) * (size_t)nsnap)) return r;
            hipLaunchKernelGGL(k_pat_snap, dim3(blocks_for(nsnap, 256)), dim3(256), 0, ctx->stream, pf, nsnap,
                               (int *)pc.snap.p, 0, d_diff);
            HIPCHK(ctx, hipGetLastError());
        }
        pc.nNodes = nNodes, pc.nElems = nElems, pc.el_stride = el_stride, pc.nCZs = nCZs;
        pc.cz_stride = cz_stride, pc.N = N, pc.cb = cb, pc.nkeys = nk;
        pc.valid = true;
    }
    // the impacts' keys, sorted and merged into the static part's
    int num = pc.nkeys;
    const unsigned long long *uniq = (const unsigned long long *)pc.keys.p;
    if (Mi > 0) {
        const long long Mt = pc.nkeys + Mi;
        if (int r = grow(ctx, ctx->pat_a, sizeof(unsigned long long) * (size_t)std::max(2 * Mi, Mt))) return r;
        if (int r = grow(ctx, ctx->pat_b, sizeof(unsigned long long) * (size_t)Mt)) return r;
        auto *a = (unsigned long long *)ctx->pat_a.p, *b = (unsigned long long *)ctx->pat_b.p;
        hipLaunchKernelGGL(k_pat_emit, dim3(blocks_for(Mi, 256)), dim3(256), 0, ctx->stream, src[2], altId, nNodes, N,
                           cb, a, d_err);
        if (int r = sort_keys(a, a + Mi, Mi)) return r;
        size_t mb = 0;
        HIPCHK(ctx, hipcub::DeviceMerge::MergeKeys(nullptr, mb, (const unsigned long long *)pc.keys.p, pc.nkeys,
                                                   (const unsigned long long *)(a + Mi), (int)Mi, b,
                                                   ::rocprim::less<unsigned long long>(), ctx->stream));
        if (int r = grow(ctx, ctx->scan_tmp, mb)) return r;
        HIPCHK(ctx, hipcub::DeviceMerge::MergeKeys(ctx->scan_tmp.p, mb, (const unsigned long long *)pc.keys.p,
                                                   pc.nkeys, (const unsigned long long *)(a + Mi), (int)Mi, b,
                                                   ::rocprim::less<unsigned long long>(), ctx->stream));
        if (int r = distinct(b, Mt, a, &num)) return r;
        uniq = a;
    }
    int err = 0;
    HIPCHK(ctx, hipMemcpyAsync(&err, d_err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (err) return fail(ctx, ICF_E_INDEX, "icf_build_pattern: a node id or altId is out of range");
    const int n_nz = num;
    *nnz = n_nz;
    if (n_nz > max_nnz) return fail(ctx, ICF_E_CAPACITY, "icf_build_pattern: more blocks than max_nnz");
    const unsigned long long *u = uniq;
    if (n_nz > 0)
        hipLaunchKernelGGL(k_pat_cols, dim3(blocks_for(n_nz, 256)), dim3(256), 0, ctx->stream, u, n_nz, cb, csr_cols);
    hipLaunchKernelGGL(k_pat_rows, dim3(blocks_for((long long)N + 1, 256)), dim3(256), 0, ctx->stream, u, n_nz, N, cb,
                       csr_rows);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int icf_pack_pcsr(icf_ctx *ctx, const int *altId, int nNodes, const int *csr_rows, const int *csr_cols, int N,
                  int *ie, int nElems, int el_stride, int *icz, int nCZs, int cz_stride, int *icr, int nImpacts,
                  int cr_stride)
{
    if (int r = check_ctx(ctx)) return r;
    if (!ctx->capturing)
        if (int r = lane_barrier(ctx)) return r;
    if (nNodes < 0 || N < 0 || bad_units(nElems, el_stride) || bad_units(nCZs, cz_stride) ||
        bad_units(nImpacts, cr_stride) || (nNodes && !altId) || !csr_rows || (nElems && !ie) || (nCZs && !icz) ||
        (nImpacts && !icr))
        return fail(ctx, ICF_E_INVALID, "icf_pack_pcsr: invalid argument");
    if (ctx->capturing) return fail(ctx, ICF_E_CAPTURE, "icf_pack_pcsr reports index errors to the host");
    int *d_err = ctx->small + 1;
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, ctx->stream, d_err, 1, 0);
    PackSrc ps[3] = {{ie, nElems, el_stride, EL_NODES, EL_SLOTS, EL_ROWS, 4},
                     {icz, nCZs, cz_stride, CZI_VRTS, CZI_SLOTS, CZI_ROWS, 6},
                     {icr, nImpacts, cr_stride, CR_NODES, CR_SLOTS, CR_ROWS, 4}};
    for (const PackSrc &p : ps) {
        if (!p.n) continue;
        hipLaunchKernelGGL(k_pack_pcsr, dim3(blocks_for((long long)p.n * p.K, 256)), dim3(256), 0, ctx->stream, p,
                           altId, nNodes, csr_rows, csr_cols, N, d_err);
        // slot and row fields changed: maps built from them are checked before their next use
        invalidate_range(ctx, (uintptr_t)(p.base + (size_t)p.stride * p.slot0),
                         (uintptr_t)(p.base + (size_t)p.stride * (p.row0 + p.K)));
    }
    if (nCZs)
        hipLaunchKernelGGL(k_pack_cz_flags, dim3(blocks_for(nCZs, 256)), dim3(256), 0, ctx->stream, icz, nCZs,
                           cz_stride);
    int err = 0;
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipMemcpyAsync(&err, d_err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (err) return fail(ctx, ICF_E_INDEX, "icf_pack_pcsr: a node id is out of range or a pair is not in the pattern");
    return 0;
}

int icf_broad_phase(icf_ctx *ctx, const double *dn, int nd_stride, const int *ie, int el_all_stride, int nTets,
                    int exclude_shared, int *pairs, long long max_pairs, long long *nOut)
{
    if (int r = check_ctx(ctx)) return r;
    if (!ctx->capturing)
        if (int r = lane_barrier(ctx)) return r;
    if (!nOut || bad_units(nTets, el_all_stride) || nd_stride < 0 || max_pairs < 0 || (nTets && (!dn || !ie)) ||
        (max_pairs && !pairs))
        return fail(ctx, ICF_E_INVALID, "icf_broad_phase: invalid argument");
    if (ctx->capturing) return fail(ctx, ICF_E_CAPTURE, "icf_broad_phase returns a host value");
    *nOut = 0;
    if (nTets == 0) return 0;
    const int nT = nTets;
    if (nT > (1 << 23)) return fail(ctx, ICF_E_CAPACITY, "icf_broad_phase: more than 2^23 surface tets");
    // scratch: lo, hi (12 nT each), bounds (7), cell-ordered k-DOPs in FP64 (24 nT) and FP32
    // (12 nT doubles' worth)
    if (int r = grow(ctx, ctx->bp_f, sizeof(double) * ((size_t)60 * nT + 8))) return r;
    if (int r = grow(ctx, ctx->bp_i, sizeof(int) * (8 * (size_t)nT + 4))) return r;
    if (int r = grow(ctx, ctx->bp_l, sizeof(long long) * BP_PARTS * BP_PART_STRIDE)) return r;
    double *lo = (double *)ctx->bp_f.p, *hi = lo + (size_t)12 * nT;
    auto *bounds = (unsigned long long *)(hi + (size_t)12 * nT);
    double2 *skd = (double2 *)(hi + (size_t)12 * nT + 8);          // 12 double2 = 24 doubles per tet
    float4 *sk32 = (float4 *)(hi + (size_t)36 * nT + 8);           // 6 float4 = 12 doubles per tet
    auto *key = (unsigned *)ctx->bp_i.p;
    int *rank = (int *)(key + nT), *sidx = rank + nT;
    int4 *snode = (int4 *)(sidx + nT + ((4 - ((3 * nT) & 3)) & 3));   // 4 nT ints, 16-byte aligned
    hipLaunchKernelGGL(k_bp_bounds_init, dim3(1), dim3(64), 0, ctx->stream, bounds);
    hipLaunchKernelGGL(k_bp_kdop, dim3(blocks_for(nT, 256)), dim3(256), 0, ctx->stream, dn, nd_stride, ie,
                       el_all_stride, nT, lo, bounds, nd_stride);
    unsigned long long hb[BP_NBOUNDS + 1];
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipMemcpyAsync(hb, bounds, sizeof(hb), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (hb[BP_NBOUNDS]) return fail(ctx, ICF_E_INDEX, "icf_broad_phase: bad node id");
    // the grid: cells at least the largest extent, coarsened (any size >= the extents finds
    // the same pairs) until the margin grid has at most 2^24 cells
    BpGrid g;
    const double ext = bp_unord(hb[0]);
    double span[3];
    for (int k = 0; k < 3; k++) {
        g.org[k] = bp_unord(hb[1 + k]);
        span[k] = bp_unord(hb[4 + k]) - g.org[k];
        if (!std::isfinite(g.org[k]) || !std::isfinite(span[k]))
            return fail(ctx, ICF_E_INVALID, "icf_broad_phase: non-finite node positions");
    }
    g.cs = ext > 0 ? ext * 1.000001 : 1.0;
    for (;;) {
        double cells = 1;
        for (int k = 0; k < 3; k++) {
            const double c = std::floor(span[k] / g.cs) + 1;
            g.n[k] = c < (double)(1 << 24) ? (int)c : (1 << 24);
            cells *= (double)g.n[k] + (k == 0 ? 3 : 2);
        }
        if (cells <= (double)(1 << 24)) break;
        g.cs *= 1.25;
    }
    g.gx = g.n[0] + 3;
    g.gy = g.n[1] + 2;
    g.gz = g.n[2] + 2;
    const int G = g.gx * g.gy * g.gz;
    // cell counts, their exclusive scan (first), and the tets' cell-ordered copies
    if (int r = grow(ctx, ctx->bp_c, sizeof(int) * 2 * ((size_t)G + 1))) return r;
    int *count = (int *)ctx->bp_c.p, *first = count + G + 1;
    size_t b3 = 0;
    HIPCHK(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, b3, count, first, G + 1, ctx->stream));
    if (int r = grow(ctx, ctx->scan_tmp, b3)) return r;
    hipLaunchKernelGGL(k_fill_int, dim3(blocks_for((long long)G + 1, 256)), dim3(256), 0, ctx->stream, count, G + 1,
                       0);
    hipLaunchKernelGGL(k_bp_cell, dim3(blocks_for(nT, 256)), dim3(256), 0, ctx->stream, lo, nT, g, key, rank, count);
    HIPCHK(ctx, hipcub::DeviceScan::ExclusiveSum(ctx->scan_tmp.p, b3, count, first, G + 1, ctx->stream));
    hipLaunchKernelGGL(k_bp_sorted, dim3(blocks_for(nT, 256)), dim3(256), 0, ctx->stream, dn, nd_stride, nd_stride,
                       nT, key, rank,
                       first, ie, el_all_stride, sidx, skd, snode, sk32);
    // candidate pairs per cell (FP32 prefilter), collected in LDS and appended in batches; the
    // exact test; the passing pairs grouped by a (a counting sort) and placed by their rank of b
    // within the group: the pairs in (a, b) order.  The launches after the search take the
    // candidate count from device memory (their grids cover the buffer), so one read-back at
    // the end serves all; a second round only when the candidates outgrew the buffer.
    const int ncell = g.n[0] * g.n[1] * g.n[2];
    auto *tot = (unsigned long long *)ctx->bp_l.p;
    // at least 64 places per region, so that every region's grid row has a workgroup
    long long cap = std::max<long long>({ctx->bp_cap, 4LL * nT, 64LL * BP_PARTS});
    size_t b5 = 0;
    HIPCHK(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, b5, (int *)nullptr, (int *)nullptr, nT + 1, ctx->stream));
    if (int r = grow(ctx, ctx->scan_tmp, b5)) return r;
    long long ncand = 0;
    int nv = 0;
    for (int pass = 0; pass < 2; pass++) {
        if (cap >= (1LL << 31)) return fail(ctx, ICF_E_CAPACITY, "icf_broad_phase: more than 2^31 candidate pairs");
        if (int r = grow(ctx, ctx->pat_a, sizeof(unsigned long long) * (size_t)cap)) return r;
        if (int r = grow(ctx, ctx->pat_b, sizeof(unsigned long long) * (size_t)cap)) return r;
        if (int r = grow(ctx, ctx->bp_g, sizeof(int) * (2 * ((size_t)nT + 1) + (size_t)cap))) return r;
        int *acount = (int *)ctx->bp_g.p, *start = acount + nT + 1, *rnk = start + nT + 1;
        auto *pa = (unsigned long long *)ctx->pat_a.p, *grp = (unsigned long long *)ctx->pat_b.p;
        hipLaunchKernelGGL(k_fill_int, dim3(blocks_for(2 * BP_PARTS * BP_PART_STRIDE, 256)), dim3(256), 0,
                           ctx->stream, (int *)tot, 2 * BP_PARTS * BP_PART_STRIDE, 0);
        hipLaunchKernelGGL(k_fill_int, dim3(blocks_for((long long)nT + 1, 256)), dim3(256), 0, ctx->stream, acount,
                           nT + 1, 0);
        hipLaunchKernelGGL(k_bp_cellwave, dim3((ncell + BPW_WAVES - 1) / BPW_WAVES), dim3(64 * BPW_WAVES), 0,
                           ctx->stream, sk32, snode, exclude_shared, g, first, tot, pa, cap);
        const dim3 gparts(blocks_for(cap / BP_PARTS, 256), BP_PARTS);   // a grid row per region
        hipLaunchKernelGGL(k_bp_exact, gparts, dim3(256), 0, ctx->stream, skd, sk32, sidx, tot,
                           cap, pa, rnk, acount);
        HIPCHK(ctx, hipcub::DeviceScan::ExclusiveSum(ctx->scan_tmp.p, b5, acount, start, nT + 1, ctx->stream));
        hipLaunchKernelGGL(k_bp_scatter, gparts, dim3(256), 0, ctx->stream, pa, rnk, tot, cap,
                           start, grp);
        if (max_pairs > 0)
            hipLaunchKernelGGL(k_bp_place, dim3(blocks_for(cap, 256)), dim3(256), 0, ctx->stream, grp, start, nT,
                               max_pairs, pairs);
        std::vector<unsigned long long> got((size_t)BP_PARTS * BP_PART_STRIDE);
        HIPCHK(ctx, hipGetLastError());
        HIPCHK(ctx, hipMemcpyAsync(got.data(), tot, sizeof(unsigned long long) * got.size(), hipMemcpyDeviceToHost,
                                   ctx->stream));
        HIPCHK(ctx, hipMemcpyAsync(&nv, start + nT, sizeof(nv), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        // the fullest region decides whether the buffer held every candidate
        ncand = 0;
        for (int q = 0; q < BP_PARTS; q++) ncand = std::max(ncand, (long long)got[(size_t)q * BP_PART_STRIDE]);
        if (ncand <= cap / BP_PARTS) break;
        if (pass == 1) return fail(ctx, ICF_E_CAPACITY, "icf_broad_phase: pair count changed between two searches");
        cap = BP_PARTS * (ncand + ncand / 4);
    }
    ctx->bp_cap = std::max(ctx->bp_cap, BP_PARTS * (ncand + ncand / 4));
    *nOut = (long long)nv;
    if ((long long)nv > max_pairs) return fail(ctx, ICF_E_CAPACITY, "icf_broad_phase: more pairs than max_pairs");
    return 0;
}

int icf_nl2set(icf_ctx *ctx, const int *masks, const int *pairs, int nPairs, const int *ie, int el_all_stride,
               int *nl2, int tet_stride, int *nOut)
{
    if (int r = check_ctx(ctx)) return r;
    if (!ctx->capturing)
        if (int r = lane_barrier(ctx)) return r;
    if (!nOut || nPairs < 0 || el_all_stride < 0 || tet_stride < 0 || (nPairs && (!masks || !pairs || !ie)) ||
        (tet_stride && !nl2))
        return fail(ctx, ICF_E_INVALID, "icf_nl2set: invalid argument");
    if (ctx->capturing) return fail(ctx, ICF_E_CAPTURE, "icf_nl2set returns a host value");
    *nOut = 0;
    if (nPairs == 0) return 0;
    if (int r = note_write(ctx, nl2, sizeof(int) * 2 * (size_t)tet_stride, "icf_nl2set")) return r;
    if (int r = grow(ctx, ctx->bp_i, sizeof(int) * 2 * ((size_t)nPairs + 1))) return r;
    int *cnt = (int *)ctx->bp_i.p, *off = cnt + nPairs + 1;
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, ctx->stream, cnt + nPairs, 1, 0);
    hipLaunchKernelGGL(k_nl2_count, dim3(blocks_for(nPairs, 256)), dim3(256), 0, ctx->stream, masks, nPairs, cnt);
    size_t b1 = 0;
    HIPCHK(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, b1, cnt, off, nPairs + 1, ctx->stream));
    if (int r = grow(ctx, ctx->scan_tmp, b1)) return r;
    HIPCHK(ctx, hipcub::DeviceScan::ExclusiveSum(ctx->scan_tmp.p, b1, cnt, off, nPairs + 1, ctx->stream));
    int total = 0;
    HIPCHK(ctx, hipMemcpyAsync(&total, off + nPairs, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (total == 0) return 0;
    if (total > (1 << 28)) return fail(ctx, ICF_E_CAPACITY, "icf_nl2set: more than 2^28 tuples");
    // the table: a power of two at least twice the tuples
    unsigned long long H = 1;
    while (H < 2ULL * (unsigned long long)total) H <<= 1;
    if (int r = grow(ctx, ctx->pat_a, sizeof(unsigned long long) * ((size_t)total + H))) return r;
    if (int r = grow(ctx, ctx->bp_l, sizeof(int) * ((size_t)H + 2 * (size_t)total + 2))) return r;
    auto *key = (unsigned long long *)ctx->pat_a.p, *tkey = key + total;
    int *tpos = (int *)ctx->bp_l.p, *keep = tpos + H, *dst = keep + total + 1;
    hipLaunchKernelGGL(k_fill_u64, dim3(blocks_for((long long)H, 256)), dim3(256), 0, ctx->stream, tkey, (long long)H,
                       NL2_EMPTY);
    hipLaunchKernelGGL(k_fill_int, dim3(blocks_for((long long)H, 256)), dim3(256), 0, ctx->stream, tpos, (int)H,
                       INT_MAX);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, ctx->stream, keep + total, 1, 0);
    hipLaunchKernelGGL(k_nl2_emit, dim3(blocks_for(nPairs, 256)), dim3(256), 0, ctx->stream, masks, pairs, nPairs, ie,
                       el_all_stride, off, key, tkey, tpos, H - 1);
    hipLaunchKernelGGL(k_nl2_keep, dim3(blocks_for(total, 256)), dim3(256), 0, ctx->stream, key, total, tkey, tpos,
                       H - 1, keep);
    size_t s1 = 0;
    HIPCHK(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, s1, keep, dst, total + 1, ctx->stream));
    if (int r = grow(ctx, ctx->scan_tmp, s1)) return r;
    HIPCHK(ctx, hipcub::DeviceScan::ExclusiveSum(ctx->scan_tmp.p, s1, keep, dst, total + 1, ctx->stream));
    // the tuples that fit are written before the count is read, so one wait serves both
    hipLaunchKernelGGL(k_nl2_write, dim3(blocks_for(total, 256)), dim3(256), 0, ctx->stream, key, total, keep, dst,
                       nl2, tet_stride);
    int num = 0;
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipMemcpyAsync(&num, dst + total, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    *nOut = num;
    if (num > tet_stride) return fail(ctx, ICF_E_CAPACITY, "icf_nl2set: more tuples than tet_stride");
    return 0;
}

}  // extern "C"
