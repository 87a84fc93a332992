// Host side of the batched parameter-scan API (cp3h_batch_* in include/cp3hip.h).
//
// A batch owns concatenated device arrays for all sets (grids, Earth models, layer lengths,
// mixed-precision spectrum tables) sized once at creation, float results allocated at the
// first unbinned computation (binned mode writes bin averages to the caller's buffer), two
// pinned arrays of per-set descriptors (hypothesis + offsets), one of which is copied to the
// device at each computation, one stream, and optionally instantiated HIP graphs of
// {descriptor copy, set-up kernel, persistent pair kernel}, one per descriptor array.  A
// parameter-scan step writes new hypotheses into the pending array while the device still
// works on the previous step from the other array, then replays that array's graph.
#include "../../include/cp3hip.h"
#include "common.hpp"
#include "host_util.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

using cp3host::fail;

struct cp3h_batch {
    int device = 0, n_sets = 0, cu_count = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_pairs = nullptr, ev_stop = nullptr;
    std::vector<cp3::BatchSet> layout;
    cp3::BatchSet* pinned = nullptr;           // the pending descriptors (= pinned_buf[next])
    cp3::BatchSet* pinned_buf[2] = {nullptr, nullptr};
    int next = 0;
    hipEvent_t ev_done[2] = {nullptr, nullptr};  // end of the last computation from each array
    cp3::BatchSet* d_sets = nullptr;
    std::vector<long long> geom_prefix, spec_prefix, item_prefix;
    long long* d_prefix = nullptr;      // geom | spec | item prefixes, (n_sets + 1) each
    double* d_cos = nullptr;
    int* d_maxlayers = nullptr;
    double* d_energies = nullptr;
    double* d_radii = nullptr;
    double* d_rhos = nullptr;
    double* d_geom = nullptr;
    float4* d_Q = nullptr;
    double2* d_R = nullptr;
    unsigned long long* d_next = nullptr;   // pair-kernel work queue
    int* d_limits = nullptr;                // per-row energy limits, parallel to the cosines
    long long cos_total = 0;
    float* d_out = nullptr;             // unbinned results, allocated on first use
    long long out_n = 0;
    float* bin_out = nullptr;           // binned mode
    int bin_nodes = 0;
    float bin_w[cp3::kMaxBinNodes] = {};
    std::vector<char> has_grid, has_density, has_params;
    int waves = 8, groups_per_cu = 4;
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;             // graph of the pending array (= execs[next])
    hipGraph_t graphs[2] = {nullptr, nullptr};
    hipGraphExec_t execs[2] = {nullptr, nullptr};
    bool launched = false, last_was_graph = false, ever_launched = false;

    cp3::BatchDesc desc() const {
        const long long n1 = n_sets + 1;
        cp3::BatchDesc d{d_sets, n_sets, d_prefix, d_prefix + n1, d_prefix + 2 * n1, d_cos, d_maxlayers,
                         d_energies, d_radii, d_rhos, d_geom, d_Q, d_R, d_next, d_limits, d_out, bin_out,
                         bin_nodes, {}};
        for (int g = 0; g < cp3::kMaxBinNodes; g++) d.bin_w[g] = bin_w[g];
        return d;
    }
    int setup_blocks() const {
        const long long cells = geom_prefix.back() + spec_prefix.back();
        const long long want = (cells + cp3::kBlock - 1) / cp3::kBlock;
        return int(std::max<long long>(1, std::min<long long>(want, 16LL * cu_count)));
    }
    int pair_blocks() const { return int(cp3::round_up(unsigned(cu_count * groups_per_cu), cp3::kXcds)); }
};

namespace {

int check(cp3h_batch* b, int set) {
    if (!b) return fail(CP3H_ERR_ARGUMENT, "null batch handle");
    if (set < 0 || set >= b->n_sets) return fail(CP3H_ERR_ARGUMENT, "set index out of range");
    return CP3H_OK;
}

int drain(cp3h_batch* b) {
    CP3H_HIP(hipSetDevice(b->device));
    CP3H_HIP(hipStreamSynchronize(b->stream));
    return CP3H_OK;
}

template <typename T>
void free_dev(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

// the unbinned result array, allocated the first time something needs it
int ensure_out(cp3h_batch* b) {
    if (b->d_out) return CP3H_OK;
    CP3H_HIP(hipSetDevice(b->device));
    CP3H_HIP(hipMalloc(&b->d_out, sizeof(float) * std::max(b->out_n, 1LL)));
    return CP3H_OK;
}

void drop_graph(cp3h_batch* b) {
    for (int k = 0; k < 2; k++) {
        if (b->execs[k]) (void)hipGraphExecDestroy(b->execs[k]);
        if (b->graphs[k]) (void)hipGraphDestroy(b->graphs[k]);
        b->execs[k] = nullptr;
        b->graphs[k] = nullptr;
    }
    b->exec = nullptr;
    b->graph = nullptr;
}

int upload(cp3h_batch* b, void* dst, const void* src, std::size_t bytes) {
    if (bytes == 0) return CP3H_OK;
    CP3H_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, b->stream));
    CP3H_HIP(hipStreamSynchronize(b->stream));
    return CP3H_OK;
}

// descriptor upload + set-up + pairs; with `timed`, an event between set-up and pairs
int enqueue(cp3h_batch* b, bool timed) {
    const cp3::BatchDesc d = b->desc();
    CP3H_HIP(hipMemcpyAsync(b->d_sets, b->pinned, sizeof(cp3::BatchSet) * b->n_sets, hipMemcpyHostToDevice,
                            b->stream));
    CP3H_HIP(cp3::launch_batch_setup(d, b->setup_blocks(), b->stream));
    if (timed) CP3H_HIP(hipEventRecord(b->ev_pairs, b->stream));
    if (b->item_prefix.back() > 0) CP3H_HIP(cp3::launch_batch_pairs(d, b->pair_blocks(), b->waves, b->stream));
    return CP3H_OK;
}

int capture(cp3h_batch* b) {
    drop_graph(b);
    CP3H_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue(b, false);
    hipGraph_t g = nullptr;
    const hipError_t end = hipStreamEndCapture(b->stream, &g);
    if (rc != CP3H_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (end != hipSuccess) return fail(CP3H_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(end));
    b->graphs[b->next] = g;
    CP3H_HIP(hipGraphInstantiate(&b->execs[b->next], g, nullptr, nullptr, 0));
    b->graph = g;
    b->exec = b->execs[b->next];
    return CP3H_OK;
}

}  // namespace

extern "C" {

int cp3h_batch_create(int device, int n_sets, const int* n_cosines, const int* n_energies, const int* n_rows,
                      cp3h_batch** out) {
    if (!out) return fail(CP3H_ERR_ARGUMENT, "null output handle pointer");
    *out = nullptr;
    if (n_sets <= 0 || !n_cosines || !n_energies || !n_rows) return fail(CP3H_ERR_ARGUMENT, "empty batch");
    for (int i = 0; i < n_sets; i++)
        if (n_cosines[i] < 0 || n_energies[i] < 0 || n_rows[i] < 1)
            return fail(CP3H_ERR_ARGUMENT, "set " + std::to_string(i) + ": bad grid or model size");
    int n_dev = 0;
    if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev == 0) return fail(CP3H_ERR_DEVICE, "No GPU found");
    if (device < 0 || device >= n_dev)
        return fail(CP3H_ERR_DEVICE, "The requested GPU Id " + std::to_string(device) + " is not available.");

    cp3h_batch* b = new cp3h_batch();
    b->device = device;
    b->n_sets = n_sets;
    b->layout.resize(n_sets);
    b->geom_prefix.assign(n_sets + 1, 0);
    b->spec_prefix.assign(n_sets + 1, 0);
    b->item_prefix.assign(n_sets + 1, 0);
    b->has_grid.assign(n_sets, 0);
    b->has_density.assign(n_sets, 0);
    b->has_params.assign(n_sets, 0);
    long long cos_n = 0, e_n = 0, model_n = 0, geom_n = 0, tab_n = 0, out_n = 0;
    for (int i = 0; i < n_sets; i++) {
        cp3::BatchSet& s = b->layout[i];
        std::memset(&s, 0, sizeof(s));
        s.n_c = n_cosines[i];
        s.n_e = n_energies[i];
        s.rows = n_rows[i];
        s.ne_pad = int(cp3::round_up(unsigned(std::max(s.n_e, 1)), cp3::kWave));
        s.cos_off = cos_n;
        s.e_off = e_n;
        s.model_off = model_n;
        s.geom_off = geom_n;
        s.tab_off = tab_n;
        s.out_row = int(cp3::round_up(unsigned(s.n_e), cp3::kOutRowAlign));
        s.out_plane = (long long)s.n_c * s.out_row;
        s.out_off = out_n;                   // a multiple of kOutRowAlign
        const long long geom = (long long)s.n_c * (s.rows + 1);
        const long long cells = (long long)(s.rows + 1) * s.ne_pad;
        cos_n += s.n_c;
        e_n += s.n_e;
        model_n += s.rows;
        geom_n += geom;
        tab_n += cells;
        out_n += 9LL * s.out_plane;
        b->geom_prefix[i + 1] = b->geom_prefix[i] + geom;
        b->spec_prefix[i + 1] = b->spec_prefix[i] + (s.n_e > 0 ? cells : 0);
        b->item_prefix[i + 1] = b->item_prefix[i] + (long long)((s.n_c + cp3::kRowsPerWave - 1) / cp3::kRowsPerWave) *
                                                    ((s.n_e + cp3::kWave - 1) / cp3::kWave);
    }
    auto bail = [&](int code) { cp3h_batch_destroy(b); return code; };
    const int rc = [&]() -> int {
        CP3H_HIP(hipSetDevice(device));
        CP3H_HIP(hipDeviceGetAttribute(&b->cu_count, hipDeviceAttributeMultiprocessorCount, device));
        CP3H_HIP(hipStreamCreate(&b->stream));
        CP3H_HIP(hipEventCreate(&b->ev_start));
        CP3H_HIP(hipEventCreate(&b->ev_pairs));
        CP3H_HIP(hipEventCreate(&b->ev_stop));
        for (int k = 0; k < 2; k++) {
            CP3H_HIP(hipHostMalloc(reinterpret_cast<void**>(&b->pinned_buf[k]), sizeof(cp3::BatchSet) * n_sets));
            std::memcpy(b->pinned_buf[k], b->layout.data(), sizeof(cp3::BatchSet) * n_sets);
            CP3H_HIP(hipEventCreateWithFlags(&b->ev_done[k], hipEventDisableTiming));
        }
        b->pinned = b->pinned_buf[0];
        CP3H_HIP(hipMalloc(&b->d_sets, sizeof(cp3::BatchSet) * n_sets));
        CP3H_HIP(hipMalloc(&b->d_prefix, sizeof(long long) * 3 * (n_sets + 1)));
        CP3H_HIP(hipMalloc(&b->d_cos, sizeof(double) * std::max(cos_n, 1LL)));
        CP3H_HIP(hipMalloc(&b->d_maxlayers, sizeof(int) * std::max(cos_n, 1LL)));
        CP3H_HIP(hipMalloc(&b->d_energies, sizeof(double) * std::max(e_n, 1LL)));
        CP3H_HIP(hipMalloc(&b->d_radii, sizeof(double) * model_n));
        CP3H_HIP(hipMalloc(&b->d_rhos, sizeof(double) * model_n));
        CP3H_HIP(hipMalloc(&b->d_geom, sizeof(double) * std::max(geom_n, 1LL)));
        CP3H_HIP(hipMalloc(&b->d_Q, sizeof(float4) * cp3::kMixedQGroups * std::max(tab_n, 1LL)));
        CP3H_HIP(hipMalloc(&b->d_R, sizeof(double2) * std::max(tab_n, 1LL)));
        CP3H_HIP(hipMalloc(&b->d_next, sizeof(unsigned long long) * cp3::kXcds * cp3::kQueueStride));
        b->cos_total = cos_n;
        b->out_n = out_n;
        std::vector<long long> prefix;
        prefix.insert(prefix.end(), b->geom_prefix.begin(), b->geom_prefix.end());
        prefix.insert(prefix.end(), b->spec_prefix.begin(), b->spec_prefix.end());
        prefix.insert(prefix.end(), b->item_prefix.begin(), b->item_prefix.end());
        return upload(b, b->d_prefix, prefix.data(), sizeof(long long) * prefix.size());
    }();
    if (rc != CP3H_OK) return bail(rc);
    *out = b;
    return CP3H_OK;
}

int cp3h_batch_destroy(cp3h_batch* b) {
    if (!b) return CP3H_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    free_dev(b->d_sets);
    free_dev(b->d_prefix);
    free_dev(b->d_cos);
    free_dev(b->d_maxlayers);
    free_dev(b->d_energies);
    free_dev(b->d_radii);
    free_dev(b->d_rhos);
    free_dev(b->d_geom);
    free_dev(b->d_Q);
    free_dev(b->d_R);
    free_dev(b->d_next);
    free_dev(b->d_limits);
    free_dev(b->d_out);
    for (int k = 0; k < 2; k++) {
        if (b->pinned_buf[k]) (void)hipHostFree(b->pinned_buf[k]);
        if (b->ev_done[k]) (void)hipEventDestroy(b->ev_done[k]);
    }
    if (b->ev_start) (void)hipEventDestroy(b->ev_start);
    if (b->ev_pairs) (void)hipEventDestroy(b->ev_pairs);
    if (b->ev_stop) (void)hipEventDestroy(b->ev_stop);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return CP3H_OK;
}

int cp3h_batch_set_grid_f64(cp3h_batch* b, int set, const double* cosines, const int* maxlayers,
                            const double* energies) {
    if (int rc = check(b, set)) return rc;
    const cp3::BatchSet& s = b->layout[set];
    if ((s.n_c > 0 && (!cosines || !maxlayers)) || (s.n_e > 0 && !energies))
        return fail(CP3H_ERR_ARGUMENT, "null grid arrays");
    for (int i = 0; i < s.n_c; i++) {
        if (!(cosines[i] >= -1.0 && cosines[i] <= 1.0)) return fail(CP3H_ERR_ARGUMENT, "cosines must lie in [-1, 1]");
        if (maxlayers[i] < 0 || maxlayers[i] > s.rows)
            return fail(CP3H_ERR_ARGUMENT, "maxlayers outside [0, rows] for set " + std::to_string(set));
    }
    for (int i = 0; i < s.n_e; i++)
        if (!(energies[i] > 0.0)) return fail(CP3H_ERR_ARGUMENT, "energies must be positive");
    if (int rc = drain(b)) return rc;
    if (int rc = upload(b, b->d_cos + s.cos_off, cosines, sizeof(double) * s.n_c)) return rc;
    if (int rc = upload(b, b->d_maxlayers + s.cos_off, maxlayers, sizeof(int) * s.n_c)) return rc;
    if (int rc = upload(b, b->d_energies + s.e_off, energies, sizeof(double) * s.n_e)) return rc;
    b->has_grid[set] = 1;
    return CP3H_OK;
}

int cp3h_batch_set_density_f64(cp3h_batch* b, int set, const double* radii_desc, const double* rhos_desc) {
    if (int rc = check(b, set)) return rc;
    const cp3::BatchSet& s = b->layout[set];
    if (!radii_desc || !rhos_desc) return fail(CP3H_ERR_ARGUMENT, "null density arrays");
    for (int i = 1; i < s.rows; i++)
        if (radii_desc[i] > radii_desc[i - 1]) return fail(CP3H_ERR_ARGUMENT, "radii must be in descending order");
    if (int rc = drain(b)) return rc;
    if (int rc = upload(b, b->d_radii + s.model_off, radii_desc, sizeof(double) * s.rows)) return rc;
    if (int rc = upload(b, b->d_rhos + s.model_off, rhos_desc, sizeof(double) * s.rows)) return rc;
    b->has_density[set] = 1;
    return CP3H_OK;
}

int cp3h_batch_set_params_f64(cp3h_batch* b, int set, const double* U18, const double* DM9, int antineutrino,
                              double production_height_cm) {
    if (int rc = check(b, set)) return rc;
    if (!U18 || !DM9) return fail(CP3H_ERR_ARGUMENT, "null parameter arrays");
    if (antineutrino != 0 && antineutrino != 1) return fail(CP3H_ERR_ARGUMENT, "neutrino type must be 0 or 1");
    if (!(production_height_cm >= 0.0)) return fail(CP3H_ERR_ARGUMENT, "production height must be >= 0");
    cp3::BatchSet& s = b->pinned[set];      // the pending array: no computation reads it
    std::memcpy(s.p.U, U18, sizeof(double) * 18);
    std::memcpy(s.p.DM, DM9, sizeof(double) * 9);
    cp3host::vacuum_order<double>(DM9, s.p.order);
    cp3host::real_basis(s.p.U, s.p.DM, s.p.Hr, s.p.Wmu);
    s.p.antinu = antineutrino;
    s.h_cm = production_height_cm;
    b->has_params[set] = 1;
    return CP3H_OK;
}

int cp3h_batch_set_params_many_f64(cp3h_batch* b, int first, int count, const double* U18s, const double* DM9s,
                                   const int* antineutrino, const double* production_height_cm) {
    if (!b) return fail(CP3H_ERR_ARGUMENT, "null batch handle");
    if (count < 0 || first < 0 || first + count > b->n_sets) return fail(CP3H_ERR_ARGUMENT, "set range out of bounds");
    if (count > 0 && (!U18s || !DM9s || !antineutrino || !production_height_cm))
        return fail(CP3H_ERR_ARGUMENT, "null parameter arrays");
    for (int i = 0; i < count; i++)
        if (int rc = cp3h_batch_set_params_f64(b, first + i, U18s + 18 * std::size_t(i), DM9s + 9 * std::size_t(i),
                                               antineutrino[i], production_height_cm[i]))
            return rc;
    return CP3H_OK;
}

int cp3h_batch_config(cp3h_batch* b, int waves_per_group, int groups_per_cu) {
    if (!b) return fail(CP3H_ERR_ARGUMENT, "null batch handle");
    if (waves_per_group != 4 && waves_per_group != 8 && waves_per_group != 16)
        return fail(CP3H_ERR_ARGUMENT, "waves per workgroup must be 4, 8 or 16");
    if (groups_per_cu < 1 || groups_per_cu > 16) return fail(CP3H_ERR_ARGUMENT, "workgroups per CU must be 1-16");
    if (int rc = drain(b)) return rc;
    if (waves_per_group != b->waves || groups_per_cu != b->groups_per_cu) drop_graph(b);
    b->waves = waves_per_group;
    b->groups_per_cu = groups_per_cu;
    return CP3H_OK;
}

int cp3h_batch_use_graph(cp3h_batch* b, int enable) {
    if (!b) return fail(CP3H_ERR_ARGUMENT, "null batch handle");
    if (int rc = drain(b)) return rc;
    b->use_graph = enable != 0;
    if (!b->use_graph) drop_graph(b);
    return CP3H_OK;
}

int cp3h_batch_compute_async(cp3h_batch* b) {
    if (!b) return fail(CP3H_ERR_ARGUMENT, "null batch handle");
    for (int i = 0; i < b->n_sets; i++)
        if (!b->has_grid[i] || !b->has_density[i] || !b->has_params[i])
            return fail(CP3H_ERR_STATE, "set " + std::to_string(i) + " is missing its grid, density or parameters");
    CP3H_HIP(hipSetDevice(b->device));
    if (b->bin_nodes == 0)
        if (int rc = ensure_out(b)) return rc;
    CP3H_HIP(hipEventRecord(b->ev_start, b->stream));
    b->exec = b->execs[b->next];
    if (b->use_graph) {
        if (!b->exec)
            if (int rc = capture(b)) return rc;
        CP3H_HIP(hipGraphLaunch(b->exec, b->stream));
    } else {
        if (int rc = enqueue(b, true)) return rc;
    }
    CP3H_HIP(hipEventRecord(b->ev_stop, b->stream));
    // the other array becomes pending once the computation that last read it has finished (the
    // one before this launch, so the device stays busy meanwhile); it starts as a copy of the
    // descriptors just launched
    CP3H_HIP(hipEventRecord(b->ev_done[b->next], b->stream));
    const int q = 1 - b->next;
    CP3H_HIP(hipEventSynchronize(b->ev_done[q]));
    std::memcpy(b->pinned_buf[q], b->pinned_buf[b->next], sizeof(cp3::BatchSet) * b->n_sets);
    b->next = q;
    b->pinned = b->pinned_buf[q];
    b->last_was_graph = b->use_graph;
    b->launched = true;
    b->ever_launched = true;
    return CP3H_OK;
}

int cp3h_batch_wait(cp3h_batch* b) {
    if (!b) return fail(CP3H_ERR_ARGUMENT, "null batch handle");
    if (int rc = drain(b)) return rc;
    b->launched = false;
    return CP3H_OK;
}

int cp3h_batch_compute(cp3h_batch* b) {
    if (int rc = cp3h_batch_compute_async(b)) return rc;
    return cp3h_batch_wait(b);
}

int cp3h_batch_last_ms(cp3h_batch* b, float* total_ms, float* pairs_ms) {
    if (!b) return fail(CP3H_ERR_ARGUMENT, "null batch handle");
    if (!b->ever_launched) return fail(CP3H_ERR_STATE, "no computation has been launched");
    CP3H_HIP(hipSetDevice(b->device));
    CP3H_HIP(hipEventSynchronize(b->ev_stop));
    if (total_ms) CP3H_HIP(hipEventElapsedTime(total_ms, b->ev_start, b->ev_stop));
    if (pairs_ms) {
        if (b->last_was_graph) *pairs_ms = -1.0f;
        else CP3H_HIP(hipEventElapsedTime(pairs_ms, b->ev_pairs, b->ev_stop));
    }
    return CP3H_OK;
}

int cp3h_batch_time_kernels(cp3h_batch* b, int reps, float* setup_ms, float* pairs_ms) {
    if (!b) return fail(CP3H_ERR_ARGUMENT, "null batch handle");
    if (reps < 1) return fail(CP3H_ERR_ARGUMENT, "reps must be >= 1");
    for (int i = 0; i < b->n_sets; i++)
        if (!b->has_grid[i] || !b->has_density[i] || !b->has_params[i])
            return fail(CP3H_ERR_STATE, "set " + std::to_string(i) + " is missing its grid, density or parameters");
    if (int rc = drain(b)) return rc;
    if (b->bin_nodes == 0)
        if (int rc = ensure_out(b)) return rc;
    const cp3::BatchDesc d = b->desc();
    CP3H_HIP(hipMemcpyAsync(b->d_sets, b->pinned, sizeof(cp3::BatchSet) * b->n_sets, hipMemcpyHostToDevice,
                            b->stream));
    CP3H_HIP(hipEventRecord(b->ev_start, b->stream));
    for (int r = 0; r < reps; r++) CP3H_HIP(cp3::launch_batch_setup(d, b->setup_blocks(), b->stream));
    CP3H_HIP(hipEventRecord(b->ev_pairs, b->stream));
    for (int r = 0; r < reps; r++) {
        CP3H_HIP(cp3::launch_reset_queues(b->d_next, b->stream));
        if (b->item_prefix.back() > 0) CP3H_HIP(cp3::launch_batch_pairs(d, b->pair_blocks(), b->waves, b->stream));
    }
    CP3H_HIP(hipEventRecord(b->ev_stop, b->stream));
    CP3H_HIP(hipEventSynchronize(b->ev_stop));
    float ms = 0.0f;
    CP3H_HIP(hipEventElapsedTime(&ms, b->ev_start, b->ev_pairs));
    if (setup_ms) *setup_ms = ms / reps;
    CP3H_HIP(hipEventElapsedTime(&ms, b->ev_pairs, b->ev_stop));
    if (pairs_ms) *pairs_ms = ms / reps;
    b->ever_launched = true;
    b->last_was_graph = true;        // cp3h_batch_last_ms: the events no longer bracket one computation
    return CP3H_OK;
}

int cp3h_batch_autotune(cp3h_batch* b, int* waves_per_group, int* groups_per_cu, float* best_ms) {
    if (!b) return fail(CP3H_ERR_ARGUMENT, "null batch handle");
    const bool graph = b->use_graph;
    b->use_graph = false;
    const int waves_opts[3] = {4, 8, 16};
    const int gpc_opts[4] = {1, 2, 4, 8};
    float best = -1.0f;
    int bw = b->waves, bg = b->groups_per_cu;
    for (int w : waves_opts)
        for (int g : gpc_opts) {
            if (w * g > 32) continue;          // at most 32 wavefronts per CU
            if (int rc = cp3h_batch_config(b, w, g)) return rc;
            float ms = 0.0f;
            for (int rep = 0; rep < 2; rep++) {    // first run warms caches and code
                if (int rc = cp3h_batch_compute(b)) return rc;
                if (int rc = cp3h_batch_last_ms(b, nullptr, &ms)) return rc;
            }
            if (best < 0.0f || ms < best) {
                best = ms;
                bw = w;
                bg = g;
            }
        }
    if (int rc = cp3h_batch_config(b, bw, bg)) return rc;
    b->use_graph = graph;
    drop_graph(b);
    if (waves_per_group) *waves_per_group = bw;
    if (groups_per_cu) *groups_per_cu = bg;
    if (best_ms) *best_ms = best;
    return CP3H_OK;
}

int cp3h_batch_copy_results_f32(cp3h_batch* b, int set, float* dst, size_t count) {
    if (int rc = check(b, set)) return rc;
    const cp3::BatchSet& s = b->layout[set];
    if (count != std::size_t(9) * s.n_c * s.n_e) return fail(CP3H_ERR_ARGUMENT, "result count must be 9*n_c*n_e");
    if (count > 0 && !dst) return fail(CP3H_ERR_ARGUMENT, "null destination");
    if (!b->d_out) return fail(CP3H_ERR_STATE, "no unbinned results have been computed");
    if (int rc = drain(b)) return rc;
    if (count)   // 9 * n_c rows of n_e floats, device rows out_row apart
        CP3H_HIP(hipMemcpy2D(dst, sizeof(float) * s.n_e, b->d_out + s.out_off, sizeof(float) * s.out_row,
                             sizeof(float) * s.n_e, std::size_t(9) * s.n_c, hipMemcpyDeviceToHost));
    return CP3H_OK;
}

int cp3h_batch_bin_average_f32(cp3h_batch* b, int first, int count, int nodes, const float* weights,
                               float* out_device) {
    if (!b) return fail(CP3H_ERR_ARGUMENT, "null batch handle");
    if (first < 0 || count < 0 || first + count > b->n_sets) return fail(CP3H_ERR_ARGUMENT, "set range out of bounds");
    if (nodes < 1 || nodes > cp3::kMaxBinNodes || !weights || (count > 0 && !out_device))
        return fail(CP3H_ERR_ARGUMENT, "bad bin-average arguments");
    if (count > 0 && !b->d_out) return fail(CP3H_ERR_STATE, "no unbinned results have been computed");
    CP3H_HIP(hipSetDevice(b->device));
    std::size_t off = 0;
    for (int i = first; i < first + count; i++) {
        const cp3::BatchSet& s = b->layout[i];
        if (s.n_e % nodes != 0) return fail(CP3H_ERR_ARGUMENT, "set energies are not a whole number of bins");
        CP3H_HIP(cp3::launch_bin_average(b->d_out + s.out_off, s.n_c, s.n_e, s.out_row, s.out_plane, nodes, weights,
                                         out_device + off, b->stream));
        off += std::size_t(9) * s.n_c * (s.n_e / nodes);
    }
    CP3H_HIP(hipStreamSynchronize(b->stream));
    return CP3H_OK;
}

int cp3h_batch_device_results(cp3h_batch* b, int set, const float** ptr) {
    if (int rc = check(b, set)) return rc;
    if (!ptr) return fail(CP3H_ERR_ARGUMENT, "null output pointer");
    if (int rc = ensure_out(b)) return rc;
    *ptr = b->d_out + b->layout[set].out_off;
    return CP3H_OK;
}

int cp3h_batch_result_strides(cp3h_batch* b, int set, long long* plane_stride, int* row_stride) {
    if (int rc = check(b, set)) return rc;
    if (!plane_stride || !row_stride) return fail(CP3H_ERR_ARGUMENT, "null output pointer");
    *plane_stride = b->layout[set].out_plane;
    *row_stride = b->layout[set].out_row;
    return CP3H_OK;
}

int cp3h_batch_set_row_limits(cp3h_batch* b, int set, const int* limits) {
    if (int rc = check(b, set)) return rc;
    const cp3::BatchSet& s = b->layout[set];
    if (limits)
        for (int i = 0; i < s.n_c; i++)
            if (limits[i] < 0 || limits[i] > s.n_e)
                return fail(CP3H_ERR_ARGUMENT, "row limits must lie in [0, n_energies]");
    if (int rc = drain(b)) return rc;
    b->launched = false;
    if (limits) {
        if (!b->d_limits) {
            CP3H_HIP(hipMalloc(&b->d_limits, sizeof(int) * std::max(b->cos_total, 1LL)));
            drop_graph(b);      // the kernel arguments now carry the limit array
        }
        if (int rc = upload(b, b->d_limits + s.cos_off, limits, sizeof(int) * s.n_c)) return rc;
    }
    b->pinned[set].has_limits = limits ? 1 : 0;
    return CP3H_OK;
}

int cp3h_batch_set_binning(cp3h_batch* b, int nodes, const float* weights, float* out_device,
                           const long long* offsets, const long long* plane_strides, const int* row_strides) {
    if (!b) return fail(CP3H_ERR_ARGUMENT, "null batch handle");
    if (nodes < 0 || nodes > cp3::kMaxBinNodes || (nodes & (nodes - 1)) != 0)
        return fail(CP3H_ERR_ARGUMENT, "nodes per bin must be 0 or a power of two up to 32");
    if (nodes > 0) {
        if (!weights || !out_device || !offsets || !plane_strides || !row_strides)
            return fail(CP3H_ERR_ARGUMENT, "null binning arrays");
        for (int i = 0; i < b->n_sets; i++) {
            const cp3::BatchSet& s = b->layout[i];
            const long long bins = s.n_e / nodes;
            if (s.n_e % nodes != 0)
                return fail(CP3H_ERR_ARGUMENT, "set " + std::to_string(i) + ": energies are not a whole number of bins");
            if (offsets[i] < 0 || row_strides[i] < bins ||
                (s.n_c > 0 && plane_strides[i] < (long long)(s.n_c - 1) * row_strides[i] + bins))
                return fail(CP3H_ERR_ARGUMENT, "set " + std::to_string(i) + ": bad output offset or strides");
        }
    }
    if (int rc = drain(b)) return rc;
    b->launched = false;
    for (int i = 0; i < b->n_sets; i++) {
        cp3::BatchSet& s = b->pinned[i];
        s.bin_off = nodes > 0 ? offsets[i] : 0;
        s.bin_plane = nodes > 0 ? plane_strides[i] : 0;
        s.bin_row = nodes > 0 ? row_strides[i] : 0;
    }
    float w[cp3::kMaxBinNodes] = {};
    for (int g = 0; g < nodes; g++) w[g] = weights[g];
    // the kernel arguments change: a captured graph must be captured again
    if (nodes != b->bin_nodes || out_device != b->bin_out || std::memcmp(w, b->bin_w, sizeof(w)) != 0) drop_graph(b);
    b->bin_nodes = nodes;
    b->bin_out = nodes > 0 ? out_device : nullptr;
    std::memcpy(b->bin_w, w, sizeof(w));
    return CP3H_OK;
}

}  // extern "C"
