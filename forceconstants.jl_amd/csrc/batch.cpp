// Batched snapshots: IFC3 assembly (LJ, mass-weighted FP32) followed by mcc3 for many
// displaced supercells of uneven size, issued as one captured HIP graph.
//
// The reference handles one system per call (src/ifc.jl:20-35, src/mcc.jl:29-51); a
// trajectory of snapshots is a loop of third_order + mcc3 calls.  Here one call takes
// the whole list:
//   plan     (host, synchronising): upload positions and masses, count neighbours per
//            snapshot, size one workspace for all snapshots' neighbour lists and pair
//            blocks plus one shared Psi buffer and one mcc3 workspace (largest snapshot).
//   tune     (optional): for every snapshot size not seen before, time each applicable
//            mcc3 configuration (kernel family x launch size) on that size's first
//            snapshot and keep the fastest; the table can be saved to and loaded from a
//            text file.
//   capture  every snapshot's kernels (neighbour fill, pair and self blocks, the dense
//            writer into the shared Psi, the mode-product GEMMs into the snapshot's K3)
//            into one graph, instantiated once and cached for the same plan.
//   launch   the graph: one launch for the whole batch.
// The graph holds no allocation, memset or host synchronisation, so replays are valid
// as long as the plan (sizes, pointers, configurations) is unchanged; a call whose plan
// differs captures a new graph.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

#include "fc3_internal.hpp"

namespace fc3 {

struct BatchState {
    void* ws = nullptr;
    size_t ws_bytes = 0;
    std::vector<char> key;          // plan the cached graph was captured for
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    size_t nodes = 0;
    int lanes = 0;                  // parallel branches of the cached graph
    // capture streams of the lanes (default priority) and their fork / join events
    std::vector<hipStream_t> lane;
    std::vector<hipEvent_t> join;
    hipEvent_t fork = nullptr;
    struct Tuned {
        Mcc3Config cfg;
        float ms = 0.0f;
    };
    std::map<std::pair<int, int>, Tuned> tuned;   // (n, precision) -> configuration
};

void release_batch(fc3_context* h)
{
    if (!h || !h->batch) return;
    BatchState* b = static_cast<BatchState*>(h->batch);
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    if (b->ws) hipFree(b->ws);
    for (hipStream_t ls : b->lane) hipStreamDestroy(ls);
    for (hipEvent_t e : b->join) hipEventDestroy(e);
    if (b->fork) hipEventDestroy(b->fork);
    delete b;
    h->batch = nullptr;
}

static BatchState* state(fc3_context* h)
{
    if (!h->batch) h->batch = new BatchState();
    return static_cast<BatchState*>(h->batch);
}

// Grow the batch workspace to at least `need` bytes (1 MiB granularity); the old buffer
// is freed first (after the stream has drained) and the cached graph, which refers to it,
// is dropped.  *moved tells whether the buffer changed (its contents are then lost).
static int grow_batch_ws(BatchState* B, size_t need, hipStream_t st, bool* moved)
{
    *moved = false;
    if (need <= B->ws_bytes) return FC3_OK;
    if (B->exec) { hipGraphExecDestroy(B->exec); B->exec = nullptr; }
    if (B->graph) { hipGraphDestroy(B->graph); B->graph = nullptr; }
    B->key.clear();
    if (B->ws) {
        FC3_HIP(hipStreamSynchronize(st));
        FC3_HIP(hipFree(B->ws));
        B->ws = nullptr;
        B->ws_bytes = 0;
    }
    const size_t want = (need + (size_t(1) << 20) - 1) & ~((size_t(1) << 20) - 1);
    hipError_t e = hipMalloc(&B->ws, want);
    if (e != hipSuccess) {
        B->ws = nullptr;
        return set_error(FC3_ERR_OUT_OF_MEMORY, "batch workspace of " + std::to_string(want) +
                                                    " bytes: " + hipGetErrorString(e));
    }
    B->ws_bytes = want;
    *moved = true;
    return FC3_OK;
}

// At least `n` lane capture streams (default priority) with their join events.
static int ensure_lane_streams(BatchState* B, int n)
{
    if (!B->fork) FC3_HIP(hipEventCreateWithFlags(&B->fork, hipEventDisableTiming));
    while ((int)B->lane.size() < n) {
        hipStream_t ls = nullptr;
        FC3_HIP(hipStreamCreateWithFlags(&ls, hipStreamNonBlocking));
        B->lane.push_back(ls);
        hipEvent_t e = nullptr;
        FC3_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        B->join.push_back(e);
    }
    return FC3_OK;
}

namespace {

struct SnapPlan {
    int N = 0, cap = 0;
    long n = 0;
    size_t off_pos = 0, off_m = 0, off_count = 0, off_nbr = 0, off_rev = 0, off_E3 = 0, off_E2 = 0,
           off_S3 = 0;
    Mcc3Config cfg;
};

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

template <class T>
void put(std::vector<char>& k, const T& v)
{
    const char* p = reinterpret_cast<const char*>(&v);
    k.insert(k.end(), p, p + sizeof(T));
}

struct Params {
    double sigma, epsilon, pot_r_cut, r_cut, tol;
    int precision;
};

// Variant 4: the fused path (no n^3 Psi: sparse mode 1 from the pair blocks, mode 2
// over each atom's neighbourhood columns, mode 3), for BF16X3 where it applies.
constexpr int kVariantFused = 4;
// Launch sizes (workgroups per GEMM launch) the tuner tries for the dense kernels; 0 is
// the library default.  A loaded table may name only these.
constexpr long kTileCands[] = {8192L, 16384L, 32768L};
bool tiles_known(long t)
{
    if (t == 0) return true;
    for (long c : kTileCands)
        if (t == c) return true;
    return false;
}

bool fused_applies(long n, int cap, int precision, const void* k3)
{
    return precision == FC3_PREC_BF16X3 && n % 4 == 0 && n >= 16 &&
           lj_mode1_sparse_lds(cap) <= 64 * 1024 && (uintptr_t)k3 % 16 == 0;
}

// Scratch of the fused path: phi^T, the gathered mode-1 result, then modes 2-3.
size_t fused_bytes(long n, int cap)
{
    return align256(sizeof(float) * (size_t)n * n) +
           align256(sizeof(uint16_t) * 2 * 3 * (size_t)n * mcc3_gather_k(cap) * (n / 3)) +
           mcc3_from_gathered_bytes((int)n, cap, true);
}

// Kernels of one snapshot on stream s (no synchronisation, no allocation).
int snapshot_work(hipStream_t s, char* base, const SnapPlan& P, const fc3_snapshot& sn,
                  const Params& prm, float* psi32, void* mws, size_t mws_bytes)
{
    NeighborList nl;
    nl.n_atoms = P.N;
    nl.cap = P.cap;
    nl.count = reinterpret_cast<int*>(base + P.off_count);
    nl.nbr = reinterpret_cast<int*>(base + P.off_nbr);
    nl.rev = reinterpret_cast<int*>(base + P.off_rev);
    const double* pos = reinterpret_cast<const double*>(base + P.off_pos);
    const double* m = sn.masses ? reinterpret_cast<const double*>(base + P.off_m) : nullptr;
    double* E3 = reinterpret_cast<double*>(base + P.off_E3);
    double* E2 = reinterpret_cast<double*>(base + P.off_E2);
    double* S3 = reinterpret_cast<double*>(base + P.off_S3);
    FC3_TRY(neighbor_fill(s, pos, P.N, sn.box, prm.r_cut, nl));
    FC3_TRY(lj_pair_blocks(s, pos, sn.box, prm.sigma, prm.epsilon, nl, E3, S3, E2, nullptr));
    Ifc3Blocks b;
    b.E = E3;
    b.S = S3;
    if (P.cfg.variant == kVariantFused) {
        const int n = (int)P.n;
        const int Kc = mcc3_gather_k(P.cap);
        char* w = static_cast<char*>(mws);
        float* phiT = reinterpret_cast<float*>(w);
        w += align256(sizeof(float) * (size_t)n * n);
        __bf16* U = reinterpret_cast<__bf16*>(w);
        w += align256(sizeof(uint16_t) * 2 * 3 * (size_t)n * Kc * P.N);
        const size_t used = (size_t)(w - static_cast<char*>(mws));
        if (used > mws_bytes) return set_error(FC3_ERR_OUT_OF_MEMORY, "batch fused scratch");
        FC3_TRY(transpose_square(s, sn.phi, n, phiT));
        FC3_TRY(lj_mode1_gathered(s, nl, b, prm.tol, m, phiT, Kc, U, nullptr, 0));
        return mcc3_from_gathered(s, U, sn.phi, phiT, n, nl.count, nl.nbr, P.cap, sn.k3, w,
                                  mws_bytes - used);
    }
    FC3_TRY(write_ifc3_pairwise(s, nl, b, prm.tol, m, nullptr, psi32));
    return mcc3_run(s, psi32, sn.phi, (int)P.n, prm.precision, sn.k3, mws, mws_bytes, &P.cfg);
}

// ---- lanes ----------------------------------------------------------------------
// The batch runs as parallel branches of one graph.  Snapshots whose K3 outputs overlap
// form a chain that stays in one lane in submission order (so a shared output buffer ends
// with the result a sequential loop would leave); the chains of one size class (largest
// n of the chain) are dealt round robin over that class's lanes.  A lane owns a Psi
// buffer and an mcc3 workspace sized for its snapshots' configurations and runs its
// snapshots one after another; the lanes run concurrently.  Lane counts shrink, the
// class with the largest per-lane need first, until the lanes fit `budget` bytes; at one
// lane per class and still too large, everything runs in a single lane.
constexpr int kMaxLanes = 16;
constexpr int kMaxLanesPerClass = 8;
// Lanes a size class can use.  A snapshot whose largest GEMM (mode 3: n^2 x n in 256 x 256
// tiles) has kLaneFillTiles tiles fills the GPU by itself, and a second lane of such
// snapshots only competes for the caches: at 1500 DoF (52.7K tiles) two lanes ran 32
// snapshots in 1126 ms against 1068 ms one after another, while at 324 DoF eight lanes
// took 10.5 ms against 12.4 ms.
constexpr long kLaneFillTiles = 32768;
int class_lane_cap(long n)
{
    const long tiles = ((n * n + 255) / 256) * ((n + 255) / 256);
    const long cap = (kLaneFillTiles + tiles - 1) / tiles;
    return (int)std::max(1L, std::min<long>(cap, kMaxLanesPerClass));
}

struct LanePlan {
    std::vector<std::vector<int>> lanes;   // snapshot indices per lane, ascending
    std::vector<size_t> off, off_mws, mws; // lane start, its workspace offset and bytes
    size_t bytes = 0;                      // the whole lane region
};

size_t snap_psi_bytes(const SnapPlan& p)
{
    return p.cfg.variant == kVariantFused ? 0 : sizeof(float) * (size_t)p.n * p.n * p.n;
}

size_t snap_mws_bytes(const SnapPlan& p, int precision)
{
    return p.cfg.variant == kVariantFused ? fused_bytes(p.n, p.cap)
                                          : mcc3_workspace_bytes((int)p.n, precision, false, &p.cfg);
}

LanePlan layout_lanes(const std::vector<std::vector<int>>& lanes, const std::vector<SnapPlan>& P,
                      int precision)
{
    LanePlan L;
    L.lanes = lanes;
    for (const auto& ln : lanes) {
        size_t psi = 0, mws = 0;
        for (int i : ln) {
            psi = std::max(psi, snap_psi_bytes(P[i]));
            mws = std::max(mws, snap_mws_bytes(P[i], precision));
        }
        L.off.push_back(L.bytes);
        L.off_mws.push_back(align256(psi));
        L.mws.push_back(mws);
        L.bytes += align256(psi) + align256(mws);
    }
    return L;
}

LanePlan plan_lanes(const std::vector<SnapPlan>& P, const fc3_snapshot* snaps, int count,
                    int precision, size_t budget)
{
    // chains: union of snapshots whose K3 byte ranges overlap
    std::vector<int> parent(count);
    for (int i = 0; i < count; ++i) parent[i] = i;
    auto find = [&](int x) {
        while (parent[x] != x) x = parent[x] = parent[parent[x]];
        return x;
    };
    std::vector<int> ord(count);
    for (int i = 0; i < count; ++i) ord[i] = i;
    std::sort(ord.begin(), ord.end(), [&](int a, int b) {
        return reinterpret_cast<uintptr_t>(snaps[a].k3) < reinterpret_cast<uintptr_t>(snaps[b].k3);
    });
    uintptr_t run_end = 0;
    int run = -1;
    for (int i : ord) {
        const uintptr_t s0 = reinterpret_cast<uintptr_t>(snaps[i].k3);
        const uintptr_t e0 = s0 + sizeof(float) * (uintptr_t)P[i].n * P[i].n * P[i].n;
        if (run >= 0 && s0 < run_end) {
            parent[find(i)] = find(run);
            run_end = std::max(run_end, e0);
        } else {
            run = i;
            run_end = e0;
        }
    }
    // chains in order of their first snapshot, grouped by class
    std::map<int, std::vector<int>> chain;   // root -> members (ascending)
    std::vector<int> first;                  // roots in order of first member
    for (int i = 0; i < count; ++i) {
        const int r = find(i);
        if (!chain.count(r)) first.push_back(r);
        chain[r].push_back(i);
    }
    std::map<long, std::vector<int>> cls;    // class n -> chain roots
    for (int r : first) {
        long n = 0;
        for (int i : chain[r]) n = std::max(n, P[i].n);
        cls[n].push_back(r);
    }
    std::map<long, int> nl;
    int total = 0;
    for (const auto& kv : cls) {
        nl[kv.first] = std::min({(int)kv.second.size(), kMaxLanesPerClass,
                                 class_lane_cap(kv.first)});
        total += nl[kv.first];
    }
    while (total > kMaxLanes) {
        auto big = std::max_element(nl.begin(), nl.end(),
                                    [](const auto& a, const auto& b) { return a.second < b.second; });
        --big->second;
        --total;
    }
    auto build = [&]() {
        std::vector<std::vector<int>> lanes;
        std::vector<long> lane_cls;
        for (const auto& kv : cls) {
            const int k = nl[kv.first];
            const size_t l0 = lanes.size();
            lanes.resize(l0 + k);
            lane_cls.resize(l0 + k, kv.first);
            for (size_t c = 0; c < kv.second.size(); ++c) {
                auto& dst = lanes[l0 + c % k];
                const auto& m = chain[kv.second[c]];
                dst.insert(dst.end(), m.begin(), m.end());
            }
            for (size_t l = l0; l < lanes.size(); ++l) std::sort(lanes[l].begin(), lanes[l].end());
        }
        return std::make_pair(lanes, lane_cls);
    };
    for (;;) {
        auto b = build();
        LanePlan L = layout_lanes(b.first, P, precision);
        if (L.bytes <= budget) return L;
        // the class whose lanes need the most shrinks by one lane
        long worst = -1;
        size_t worst_need = 0;
        for (size_t l = 0; l < L.lanes.size(); ++l) {
            const size_t need = L.off_mws[l] + align256(L.mws[l]);
            if (nl[b.second[l]] > 1 && need > worst_need) {
                worst_need = need;
                worst = b.second[l];
            }
        }
        if (worst < 0) break;
        --nl[worst];
    }
    std::vector<int> all(count);
    for (int i = 0; i < count; ++i) all[i] = i;
    return layout_lanes({all}, P, precision);
}

// Device modes of every snapshot (FC3_BATCH_DEVICE_MODES): get_modes(dynamical_matrix(sys,
// LJ, calc)) of src/dynmat.jl:5-33,123-131 on the device, the eigenvectors written to the
// snapshot's phi in FP32 and the eigenvalues to its freqs_sq (host, if given) with the 3
// of smallest magnitude set to zero.  IFC2 is built from the snapshot's neighbour list and
// pair blocks in the batch workspace; the snapshots of one size share rocSOLVER
// dsyevd_strided_batched calls (chunks of at most 4 GiB of matrices), which take the place
// of one latency-bound eigensolver call per snapshot.  Runs before the graph, outside it.
constexpr size_t kModesChunkBytes = size_t(4) << 30;
constexpr int kRigidModes = 3;   // get_modes' default num_rigid_translation

int batch_device_modes(fc3_context* h, char* base, const std::vector<SnapPlan>& P,
                       const fc3_snapshot* snaps, int count, const Params& prm)
{
    hipStream_t s = h->stream;
    std::map<long, std::vector<int>> by_n;
    for (int i = 0; i < count; ++i) by_n[P[i].n].push_back(i);
    for (const auto& kv : by_n) {
        const int n = (int)kv.first, Nat = n / 3;
        const std::vector<int>& idx = kv.second;
        const size_t per = sizeof(double) * ((size_t)n * n + 2 * (size_t)n) + 64;
        const size_t G = std::max<size_t>(1, std::min<size_t>(idx.size(), kModesChunkBytes / per));
        for (size_t c0 = 0; c0 < idx.size(); c0 += G) {
            const int g = (int)std::min<size_t>(G, idx.size() - c0);
            size_t sv = 0;
            FC3_TRY(dsyevd_batched_workspace_bytes(h, n, g, &sv));
            const size_t need = Carve::need(sizeof(double) * (size_t)n * n * g) +
                                2 * Carve::need(sizeof(double) * (size_t)n * g) +
                                Carve::need(sizeof(int) * g) +
                                Carve::need(sizeof(double) * 9 * (size_t)Nat) + Carve::need(sv) +
                                4096;
            FC3_TRY(ensure_workspace(h, need));
            Carve cv(h->ws, h->ws_bytes);
            double* A = cv.take<double>((size_t)n * n * g);
            double* W = cv.take<double>((size_t)n * g);
            double* E = cv.take<double>((size_t)n * g);
            int* info = cv.take<int>(g);
            double* S2 = cv.take<double>(9 * (size_t)Nat);
            char* svw = cv.take<char>(sv);
            if (!cv.ok()) return set_error(FC3_ERR_OUT_OF_MEMORY, "batch modes workspace");
            for (int t = 0; t < g; ++t) {
                const int i = idx[c0 + t];
                const SnapPlan& p = P[i];
                NeighborList nl;
                nl.n_atoms = p.N;
                nl.cap = p.cap;
                nl.count = reinterpret_cast<int*>(base + p.off_count);
                nl.nbr = reinterpret_cast<int*>(base + p.off_nbr);
                nl.rev = reinterpret_cast<int*>(base + p.off_rev);
                const double* pos = reinterpret_cast<const double*>(base + p.off_pos);
                const double* m = reinterpret_cast<const double*>(base + p.off_m);
                double* E3 = reinterpret_cast<double*>(base + p.off_E3);
                double* E2 = reinterpret_cast<double*>(base + p.off_E2);
                double* Ai = A + (size_t)t * n * n;
                FC3_TRY(neighbor_fill(s, pos, p.N, snaps[i].box, prm.r_cut, nl));
                FC3_TRY(lj_pair_blocks(s, pos, snaps[i].box, prm.sigma, prm.epsilon, nl, E3,
                                       nullptr, E2, S2));
                FC3_TRY(lj_write_ifc2(s, nl, E2, S2, prm.tol, Ai));
                FC3_TRY(dynmat_scale(s, Ai, n, m));
            }
            FC3_TRY(eigen_strided_batched(h, A, n, g, W, E, info, svw, sv));
            for (int t = 0; t < g; ++t)
                FC3_TRY(to_f32(s, A + (size_t)t * n * n, (long)n * n, snaps[idx[c0 + t]].phi));
            std::vector<double> hw((size_t)n * g);
            std::vector<int> hinfo(g);
            FC3_HIP(hipMemcpyAsync(hw.data(), W, sizeof(double) * hw.size(), hipMemcpyDeviceToHost, s));
            FC3_HIP(hipMemcpyAsync(hinfo.data(), info, sizeof(int) * g, hipMemcpyDeviceToHost, s));
            FC3_HIP(hipStreamSynchronize(s));
            for (int t = 0; t < g; ++t) {
                const int i = idx[c0 + t];
                if (hinfo[t] != 0)
                    return set_error(FC3_ERR_HIP, "snapshot " + std::to_string(i) +
                                                      ": rocSOLVER dsyevd did not converge (info " +
                                                      std::to_string(hinfo[t]) + ")");
                if (!snaps[i].freqs_sq) continue;
                double* w = snaps[i].freqs_sq;
                std::copy(hw.begin() + (size_t)t * n, hw.begin() + (size_t)(t + 1) * n, w);
                // get_modes: the eigenvalues of smallest magnitude (stable order, as
                // sortperm) are set to zero; the vectors are untouched
                std::vector<int> ord(n);
                for (int a = 0; a < n; ++a) ord[a] = a;
                std::stable_sort(ord.begin(), ord.end(),
                                 [&](int a, int b) { return std::fabs(w[a]) < std::fabs(w[b]); });
                for (int r = 0; r < std::min(kRigidModes, n); ++r) w[ord[r]] = 0.0;
            }
        }
    }
    return FC3_OK;
}

}  // namespace

static int tune_size(fc3_context* h, BatchState* B, char* base, const SnapPlan& P,
                     const fc3_snapshot& sn, const Params& prm, float* psi32, void* mws,
                     size_t mws_bytes)
{
    // candidates of the requested precision: the fused path and the pre-split kernels at
    // three launch sizes for BF16X3, else (FP32, or BF16X3 with n % 4 != 0) the FP32
    // kernels; each is timed over the whole snapshot (assembly included) after a warm-up
    std::vector<Mcc3Config> cands;
    if (fused_applies(P.n, P.cap, prm.precision, sn.k3))
        cands.push_back(Mcc3Config{kVariantFused, 0});
    if (mcc3_variant_applies((int)P.n, prm.precision, 3))
        for (long t : kTileCands) cands.push_back(Mcc3Config{3, t});
    else
        cands.push_back(Mcc3Config{2, 0});
    BatchState::Tuned best;
    best.ms = -1.0f;
    for (const Mcc3Config& c : cands) {
        if (c.variant != kVariantFused &&
            mcc3_workspace_bytes((int)P.n, prm.precision, false, &c) > mws_bytes)
            continue;
        SnapPlan Q = P;
        Q.cfg = c;
        FC3_TRY(snapshot_work(h->stream, base, Q, sn, prm, psi32, mws, mws_bytes));
        FC3_HIP(hipEventRecord(h->ev_start, h->stream));
        FC3_TRY(snapshot_work(h->stream, base, Q, sn, prm, psi32, mws, mws_bytes));
        FC3_HIP(hipEventRecord(h->ev_stop, h->stream));
        FC3_HIP(hipEventSynchronize(h->ev_stop));
        float ms = 0.0f;
        FC3_HIP(hipEventElapsedTime(&ms, h->ev_start, h->ev_stop));
        if (best.ms < 0.0f || ms < best.ms) {
            best.ms = ms;
            best.cfg = c;
        }
    }
    B->tuned[{(int)P.n, prm.precision}] = best;
    return FC3_OK;
}

}  // namespace fc3

using namespace fc3;

extern "C" {

int fc3_batch_third_order_mcc3_lj(fc3_handle h, const fc3_snapshot* snaps, int count,
                                  double sigma, double epsilon, double pot_r_cut, double r_cut,
                                  double tol, int precision, int flags)
{
    FC3_TRACE_RANGE();
    if (!h) return set_error(FC3_ERR_ARGUMENT, "null handle");
    FC3_HIP(hipSetDevice(h->device));
    set_current(h);
    if (!snaps || count < 1) return set_error(FC3_ERR_ARGUMENT, "snaps must hold count >= 1 entries");
    if (precision != FC3_PREC_FP32 && precision != FC3_PREC_BF16X3)
        return set_error(FC3_ERR_ARGUMENT, "unknown precision");
    const Params prm{sigma, epsilon, pot_r_cut, r_cut, tol, precision};
    for (int i = 0; i < count; ++i) {
        const fc3_snapshot& s = snaps[i];
        if (!s.positions || !s.phi || !s.k3 || s.n_atoms < 1)
            return set_error(FC3_ERR_ARGUMENT, "snapshot " + std::to_string(i) +
                                                   ": positions, phi, k3 must be non-null, n_atoms >= 1");
        if (3LL * s.n_atoms > 65535) return set_error(FC3_ERR_UNSUPPORTED, "3N must be <= 65535");
        for (int c = 0; c < 3; ++c)
            if (!(pot_r_cut < s.box[c]))
                return set_error(FC3_ERR_CUTOFF, "Cutoff larger than L/2");
        if (s.masses)
            for (int a = 0; a < s.n_atoms; ++a)
                if (!(s.masses[a] > 0.0)) return set_error(FC3_ERR_ARGUMENT, "masses must be positive");
        if ((flags & FC3_BATCH_DEVICE_MODES) && !s.masses)
            return set_error(FC3_ERR_ARGUMENT, "snapshot " + std::to_string(i) +
                                                   ": device modes need the masses");
    }
    BatchState* B = state(h);
    hipStream_t st = h->stream;

    // ---- plan: layout of the per-snapshot region (positions, masses, counts first) ----
    std::vector<SnapPlan> P(count);
    size_t head = 0;
    long max_n = 0;
    for (int i = 0; i < count; ++i) {
        SnapPlan& p = P[i];
        p.N = snaps[i].n_atoms;
        p.n = 3L * p.N;
        max_n = std::max(max_n, p.n);
        p.off_pos = head;
        head = align256(head + sizeof(double) * 3 * p.N);
        p.off_m = head;
        head = align256(head + sizeof(double) * p.N);
        p.off_count = head;
        head = align256(head + sizeof(int) * (p.N + 1));
    }
    // neighbour counts in the handle's workspace (positions uploaded there first)
    FC3_TRY(ensure_workspace(h, head + 256));
    {
        char* w = static_cast<char*>(h->ws);
        for (int i = 0; i < count; ++i)
            FC3_HIP(hipMemcpyAsync(w + P[i].off_pos, snaps[i].positions,
                                   sizeof(double) * 3 * P[i].N, hipMemcpyHostToDevice, st));
        for (int i = 0; i < count; ++i) {
            int hmax = 0;
            int* cnt = reinterpret_cast<int*>(w + P[i].off_count);
            FC3_TRY(neighbor_count(st, reinterpret_cast<double*>(w + P[i].off_pos), P[i].N,
                                   snaps[i].box, r_cut, cnt, cnt + P[i].N, &hmax));
            P[i].cap = hmax > 0 ? hmax : 1;
        }
    }
    size_t off = head;
    size_t max_mws = 0;
    for (int i = 0; i < count; ++i) {
        SnapPlan& p = P[i];
        const size_t slots = (size_t)p.N * p.cap;
        p.off_nbr = off;
        off = align256(off + sizeof(int) * slots);
        p.off_rev = off;
        off = align256(off + sizeof(int) * slots);
        p.off_E3 = off;
        off = align256(off + sizeof(double) * 27 * slots);
        p.off_E2 = off;
        off = align256(off + sizeof(double) * 9 * slots);
        p.off_S3 = off;
        off = align256(off + sizeof(double) * 27 * p.N);
    }
    // configurations: tuned ones where known (tuning needs the buffers, so it runs below)
    const bool tune = flags & FC3_BATCH_AUTOTUNE;
    for (int i = 0; i < count; ++i) {
        auto it = B->tuned.find({(int)P[i].n, precision});
        if (it != B->tuned.end()) P[i].cfg = it->second.cfg;
        // size the mcc3 workspace for every candidate a tuning run may try
        for (int v : {2, 3})
            for (long t : kTileCands) {
                Mcc3Config c{v, t};
                if (!mcc3_variant_applies((int)P[i].n, precision, v)) continue;
                max_mws = std::max(max_mws, mcc3_workspace_bytes((int)P[i].n, precision, false, &c));
            }
        // and for the configuration this snapshot will actually run
        if (P[i].cfg.variant != kVariantFused)
            max_mws = std::max(max_mws,
                               mcc3_workspace_bytes((int)P[i].n, precision, false, &P[i].cfg));
        if (fused_applies(P[i].n, P[i].cap, precision, snaps[i].k3))
            max_mws = std::max(max_mws, fused_bytes(P[i].n, P[i].cap));
        else if (P[i].cfg.variant == kVariantFused)
            P[i].cfg = Mcc3Config{};   // a loaded table may name it where it does not apply
    }
    // tuning scratch (one shared Psi of the largest size and an mcc3 workspace for every
    // candidate), only while a size still has to be tuned; the lanes reuse these bytes
    bool need_tune = false;
    if (tune)
        for (int i = 0; i < count; ++i)
            if (!B->tuned.count({(int)P[i].n, precision})) need_tune = true;
    const size_t off_psi = off;
    const size_t off_mws = align256(off_psi + sizeof(float) * (size_t)max_n * max_n * max_n);
    const size_t regions = off;
    if (need_tune) off = align256(off_mws + max_mws);
    bool moved = false;
    FC3_TRY(grow_batch_ws(B, off, st, &moved));
    char* base = static_cast<char*>(B->ws);
    // inputs and neighbour counts into the batch workspace (outside the graph)
    auto upload = [&]() {
        for (int i = 0; i < count; ++i) {
            const SnapPlan& p = P[i];
            FC3_HIP(hipMemcpyAsync(base + p.off_pos, snaps[i].positions, sizeof(double) * 3 * p.N,
                                   hipMemcpyHostToDevice, st));
            if (snaps[i].masses)
                FC3_HIP(hipMemcpyAsync(base + p.off_m, snaps[i].masses, sizeof(double) * p.N,
                                       hipMemcpyHostToDevice, st));
            int hmax = 0;
            int* cnt = reinterpret_cast<int*>(base + p.off_count);
            FC3_TRY(neighbor_count(st, reinterpret_cast<double*>(base + p.off_pos), p.N,
                                   snaps[i].box, r_cut, cnt, cnt + p.N, &hmax));
            if (hmax > p.cap)
                return set_error(FC3_ERR_CAPACITY, "neighbour count changed between passes");
        }
        return (int)FC3_OK;
    };
    FC3_TRY(upload());

    // ---- phi on the device (before tuning, which runs the snapshots) ----
    if (flags & FC3_BATCH_DEVICE_MODES) {
        const bool prof_on = h->prof.enabled;
        h->prof.enabled = false;
        const int rc = batch_device_modes(h, base, P, snaps, count, prm);
        h->prof.enabled = prof_on;
        FC3_TRY(rc);
    }

    // ---- tune sizes not seen before ----
    const bool prof_was = h->prof.enabled;
    if (need_tune) {
        float* tpsi = reinterpret_cast<float*>(base + off_psi);
        void* tmws = base + off_mws;
        h->prof.enabled = false;
        for (int i = 0; i < count; ++i) {
            if (B->tuned.count({(int)P[i].n, precision})) continue;
            FC3_TRY(tune_size(h, B, base, P[i], snaps[i], prm, tpsi, tmws, max_mws));
        }
        h->prof.enabled = prof_was;
    }
    if (tune) {
        for (int i = 0; i < count; ++i) {
            P[i].cfg = B->tuned[{(int)P[i].n, precision}].cfg;
            if (P[i].cfg.variant == kVariantFused &&
                !fused_applies(P[i].n, P[i].cap, precision, snaps[i].k3))
                P[i].cfg = Mcc3Config{};
        }
    }

    // ---- lanes: parallel branches of the graph, each with its own Psi and scratch ----
    size_t free_b = 0, total_b = 0;
    FC3_HIP(hipMemGetInfo(&free_b, &total_b));
    const size_t reusable = B->ws_bytes - regions;   // bytes past the snapshot regions
    const size_t reserve = size_t(2) << 30;
    size_t budget = free_b + reusable > reserve ? free_b + reusable - reserve : 0;
    // FC3_BATCH_LANE_BYTES caps the lanes' bytes (to leave memory to the caller, or to
    // force fewer lanes)
    if (const char* e = std::getenv("FC3_BATCH_LANE_BYTES")) {
        const double cap = std::atof(e);
        if (cap >= 0.0 && cap < (double)budget) budget = (size_t)cap;
    }
    const LanePlan L = plan_lanes(P, snaps, count, precision, budget);
    FC3_TRY(grow_batch_ws(B, regions + L.bytes, st, &moved));
    base = static_cast<char*>(B->ws);
    if (moved) FC3_TRY(upload());   // a new buffer: positions, masses and counts again
    FC3_TRY(ensure_lane_streams(B, (int)L.lanes.size()));
    auto lane_work = [&](hipStream_t s, int l, int i) {
        char* lb = base + regions + L.off[l];
        return snapshot_work(s, base, P[i], snaps[i], prm,
                             reinterpret_cast<float*>(lb), lb + L.off_mws[l], L.mws[l]);
    };

    if (flags & FC3_BATCH_NO_GRAPH) {
        for (size_t l = 0; l < L.lanes.size(); ++l)
            for (int i : L.lanes[l]) FC3_TRY(lane_work(st, (int)l, i));
        return FC3_OK;
    }

    // ---- graph: reuse when the plan is unchanged, else capture ----
    std::vector<char> key;
    put(key, B->ws);
    put(key, prm.sigma);
    put(key, prm.epsilon);
    put(key, prm.pot_r_cut);
    put(key, prm.r_cut);
    put(key, prm.tol);
    put(key, prm.precision);
    put(key, count);
    for (int i = 0; i < count; ++i) {
        put(key, P[i].N);
        put(key, P[i].cap);
        put(key, P[i].cfg.variant);
        put(key, P[i].cfg.tiles);
        put(key, snaps[i].phi);
        put(key, snaps[i].k3);
        put(key, snaps[i].masses != nullptr);
        put(key, snaps[i].box);
    }
    put(key, regions);
    for (size_t l = 0; l < L.lanes.size(); ++l) {
        put(key, L.off[l]);
        put(key, L.off_mws[l]);
        put(key, L.mws[l]);
        put(key, L.lanes[l].size());
        for (int i : L.lanes[l]) put(key, i);
    }
    if (!B->exec || key != B->key) {
        if (B->exec) { hipGraphExecDestroy(B->exec); B->exec = nullptr; }
        if (B->graph) { hipGraphDestroy(B->graph); B->graph = nullptr; }
        B->key.clear();
        h->prof.enabled = false;   // no event records inside the capture
        FC3_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
        int rc = FC3_OK;
        hipError_t fe = hipSuccess;
        if (L.lanes.size() == 1) {
            for (int i : L.lanes[0])
                if (rc == FC3_OK) rc = lane_work(st, 0, i);
        } else {
            // fork: every lane waits for the capture origin; join: the origin waits for
            // every lane (also after an error, so that the capture can end)
            if (fe == hipSuccess) fe = hipEventRecord(B->fork, st);
            for (size_t l = 0; l < L.lanes.size(); ++l) {
                hipStream_t ls = B->lane[l];
                if (fe == hipSuccess) fe = hipStreamWaitEvent(ls, B->fork, 0);
                for (int i : L.lanes[l])
                    if (rc == FC3_OK && fe == hipSuccess) rc = lane_work(ls, (int)l, i);
                if (fe == hipSuccess) fe = hipEventRecord(B->join[l], ls);
                if (fe == hipSuccess) fe = hipStreamWaitEvent(st, B->join[l], 0);
            }
        }
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &g);
        h->prof.enabled = prof_was;
        if (rc != FC3_OK) { if (g) hipGraphDestroy(g); return rc; }
        if (fe != hipSuccess) {
            if (g) hipGraphDestroy(g);
            return set_error(FC3_ERR_HIP, std::string("graph fork/join: ") + hipGetErrorString(fe));
        }
        if (e != hipSuccess) return set_error(FC3_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        B->graph = g;
        FC3_HIP(hipGraphGetNodes(g, nullptr, &B->nodes));
        FC3_HIP(hipGraphInstantiate(&B->exec, g, nullptr, nullptr, 0));
        B->key = key;
        B->lanes = (int)L.lanes.size();
    }
    FC3_HIP(hipGraphLaunch(B->exec, st));
    return FC3_OK;
}

int fc3_batch_graph_nodes(fc3_handle h, size_t* nodes)
{
    if (!h || !nodes) return set_error(FC3_ERR_ARGUMENT, "null argument");
    *nodes = h->batch ? static_cast<BatchState*>(h->batch)->nodes : 0;
    return FC3_OK;
}

int fc3_batch_graph_lanes(fc3_handle h, int* lanes)
{
    if (!h || !lanes) return set_error(FC3_ERR_ARGUMENT, "null argument");
    *lanes = h->batch ? static_cast<BatchState*>(h->batch)->lanes : 0;
    return FC3_OK;
}

int fc3_autotune_query(fc3_handle h, int n, int precision, int* variant, long* tiles, float* ms)
{
    if (!h || !variant || !tiles || !ms) return set_error(FC3_ERR_ARGUMENT, "null argument");
    BatchState* B = state(h);
    auto it = B->tuned.find({n, precision});
    if (it == B->tuned.end()) return set_error(FC3_ERR_ARGUMENT, "size not tuned");
    *variant = it->second.cfg.variant;
    *tiles = it->second.cfg.tiles;
    *ms = it->second.ms;
    return FC3_OK;
}

int fc3_autotune_save(fc3_handle h, const char* path)
{
    if (!h || !path) return set_error(FC3_ERR_ARGUMENT, "null argument");
    BatchState* B = state(h);
    FILE* f = std::fopen(path, "w");
    if (!f) return set_error(FC3_ERR_ARGUMENT, std::string("cannot write ") + path);
    std::fprintf(f, "# n precision variant tiles ms\n");
    for (const auto& kv : B->tuned)
        std::fprintf(f, "%d %d %d %ld %.6f\n", kv.first.first, kv.first.second,
                     kv.second.cfg.variant, kv.second.cfg.tiles, kv.second.ms);
    std::fclose(f);
    return FC3_OK;
}

int fc3_autotune_load(fc3_handle h, const char* path)
{
    if (!h || !path) return set_error(FC3_ERR_ARGUMENT, "null argument");
    BatchState* B = state(h);
    FILE* f = std::fopen(path, "r");
    if (!f) return set_error(FC3_ERR_ARGUMENT, std::string("cannot read ") + path);
    char line[256];
    while (std::fgets(line, sizeof line, f)) {
        if (line[0] == '#') continue;
        int n, prec, var;
        long tiles;
        float ms;
        if (std::sscanf(line, "%d %d %d %ld %f", &n, &prec, &var, &tiles, &ms) != 5) continue;
        const bool fused_ok = var == kVariantFused && prec == FC3_PREC_BF16X3 && n % 4 == 0 &&
                              n >= 16;
        if (!fused_ok && !mcc3_variant_applies(n, prec, var)) continue;
        if (!tiles_known(tiles)) {
            std::fclose(f);
            return set_error(FC3_ERR_ARGUMENT, std::string(path) + ": tiles " +
                                                   std::to_string(tiles) +
                                                   " is not a launch size the tuner uses");
        }
        BatchState::Tuned t;
        t.cfg = Mcc3Config{var, tiles};
        t.ms = ms;
        B->tuned[{n, prec}] = t;
    }
    std::fclose(f);
    return FC3_OK;
}

}  // extern "C"
