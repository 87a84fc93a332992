// Persistent pair kernel of a batched parameter scan: many parameter sets, each with its
// own uneven grid and Earth model, propagated in one launch.
//
// In binned mode a set's energies are the quadrature nodes of its energy bins, bin-major, with
// a power-of-two node count that divides 64: each aligned group of `nodes` lanes holds one
// bin, reduces its weighted probabilities with butterfly lane exchanges, and its first lane
// stores the nine bin averages (16x less output than the unbinned results).
//
// Work items are (set, 64-energy tile, pair of cosine rows), numbered set-major, then
// tile-major, so consecutive items share a spectrum-table slice.  A fixed grid of workgroups (sized from
// the CU count by the autotuner) drains the item list through work queues, one per XCD:
// the list is cut into kXcds contiguous parts, and the wavefronts of XCD x (workgroups are
// dispatched to XCDs round robin, so x = blockIdx % kXcds) take kGrab consecutive items at a
// time from part x with one atomic add on its counter; once part x is exhausted they help
// with the other parts in turn.  So the items in flight on one XCD stay inside a narrow
// window of its own part (a few table slices, resident in that XCD's L2) however much the
// per-row cost varies (1 to 2 x shells layers per path), and the atomics spread over eight
// counters.  (A static stride pins each wavefront to one cosine row when the stride is a
// multiple of the row count, and the fast rows run ahead through the whole table; one
// global queue needs large grabs to keep its atomic rate down, which widens the window past
// L2 when sets are small.)  The counters are reset by the set-up kernel that precedes every
// launch.  Each wavefront is independent: no barriers, and every wavefront leaves once a
// grab from each of the kXcds parts has come back past that part's end.
#include "common.hpp"
#include "pair_mixed.hpp"

#include <algorithm>

namespace cp3 {
namespace {

__device__ inline int item_set(const long long* __restrict__ off, int n, long long t) {
    int lo = 0, hi = n;          // off[lo] <= t < off[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= t) lo = mid; else hi = mid;
    }
    return lo;
}

constexpr int kGrab = 4;

// lane 0's value, held in scalar registers
__device__ inline long long first_lane(unsigned long long v) {
    const unsigned lo = unsigned(__builtin_amdgcn_readfirstlane(int(unsigned(v))));
    const unsigned hi = unsigned(__builtin_amdgcn_readfirstlane(int(unsigned(v >> 32))));
    return (long long)(((unsigned long long)hi << 32) | lo);
}

// Stores one row's results: the nine P at this lane's energy, or, in binned mode, the nine bin
// averages of each aligned group of `nodes` lanes (every lane takes part in the exchanges).
template <bool kBinned>
__device__ inline void store_row(const BatchDesc& b, const BatchSet& set, int ic, int n_e, int ie, int lane,
                                 const float (&p)[3][3]) {
    if (kBinned) {
        // groups of `nodes` lanes are whole bins, all inside or all beyond the set's n_e (a row
        // limit inside a bin still leaves every node of that bin computed: the tile is whole)
        const int nodes = b.bin_nodes;
        const float w = b.bin_w[lane & (nodes - 1)];
        float v[9];
#pragma unroll
        for (int t = 0; t < 9; t++) v[t] = w * p[t / 3][t % 3];
        for (int off = nodes >> 1; off > 0; off >>= 1)
#pragma unroll
            for (int t = 0; t < 9; t++) v[t] += __shfl_xor(v[t], off);
        if ((lane & (nodes - 1)) == 0 && ie < n_e) {
            float* o = b.bin_out + set.bin_off + (long long)ic * set.bin_row + ie / nodes;
#pragma unroll
            for (int t = 0; t < 9; t++) o[t * set.bin_plane] = v[t];
        }
    } else if (ie < n_e) {
        // non-temporal: the results stream past L2 instead of evicting the table slices that the
        // other rows of the tile still read
        const std::size_t plane = set.out_plane;
        float* o = b.out + set.out_off + std::size_t(ic) * set.out_row + ie;
#pragma unroll
        for (int in = 0; in < 3; in++)
#pragma unroll
            for (int ou = 0; ou < 3; ou++) __builtin_nontemporal_store(p[in][ou], o + (in * 3 + ou) * plane);
    }
}

// One work item: 64 energies (one per lane) of a pair of cosine rows (rows 2j, 2j+1; local =
// tile * pairs + j) of set k.  Two rows that both reach this tile and cross the same number of
// shells are propagated together and share every table read (set tables are larger than L1,
// so a row on its own re-reads its slice from L2); otherwise each on its own.
template <bool kBinned>
__device__ inline void pair_item(const BatchDesc& b, int k, long long local, int lane) {
    const BatchSet& set = b.sets[k];
    const int pairs = (set.n_c + kRowsPerWave - 1) / kRowsPerWave;
    const int tile = int(local / pairs);
    const int ic0 = int(local % pairs) * kRowsPerWave;
    const int ie = tile * kWave + lane;            // < ne_pad
    auto row_ne = [&](int ic) {                    // the row's energies (a row limit ends them early)
        return set.has_limits ? min(set.n_e, b.row_limit[set.cos_off + ic]) : set.n_e;
    };
    const int ne0 = row_ne(ic0);
    const int ne1 = ic0 + 1 < set.n_c ? row_ne(ic0 + 1) : 0;
    const bool v0 = tile * kWave < ne0, v1 = tile * kWave < ne1;
    const std::size_t gstride = set.rows + 1;
    const double* gbase = b.geom + set.geom_off;
    const float4* Q = b.Q + set.tab_off * kMixedQGroups;
    const float4 wmu = make_float4(float(set.p.Wmu[0]), float(set.p.Wmu[1]), float(set.p.Wmu[2]), float(set.p.Wmu[3]));
    const double2* R = b.R + set.tab_off;
    const int L0 = v0 ? b.maxlayers[set.cos_off + ic0] : -1;
    const int L1 = v1 ? b.maxlayers[set.cos_off + ic0 + 1] : -2;
    if (L0 == L1) {
        const double* gl[2] = {gbase + std::size_t(ic0) * gstride, gbase + std::size_t(ic0 + 1) * gstride};
        float p[2][3][3];
        propagate_rows_mixed<2, kFormEigenLocal>(L0, gl, Q, nullptr, R, set.ne_pad, ie, wmu, nullptr, p);
        store_row<kBinned>(b, set, ic0, ne0, ie, lane, p[0]);
        store_row<kBinned>(b, set, ic0 + 1, ne1, ie, lane, p[1]);
        return;
    }
#pragma unroll 1
    for (int r = 0; r < 2; r++) {
        if (!(r == 0 ? v0 : v1)) continue;
        const int ic = ic0 + r;
        const double* gl[1] = {gbase + std::size_t(ic) * gstride};
        float p[1][3][3];
        propagate_rows_mixed<1, kFormEigenLocal>(r == 0 ? L0 : L1, gl, Q, nullptr, R, set.ne_pad, ie, wmu, nullptr, p);
        store_row<kBinned>(b, set, ic, r == 0 ? ne0 : ne1, ie, lane, p[0]);
    }
}

template <int kWaves, bool kBinned>
__global__ __launch_bounds__(kWaves * kWave) void batch_pairs(BatchDesc b) {
    const long long total = b.item_prefix[b.n_sets];
    const int lane = int(threadIdx.x) % kWave;
    const int home = int(blockIdx.x % kXcds);
    for (int p = 0; p < kXcds; p++) {
        const int part = (home + p) % kXcds;
        const long long part_begin = total * part / kXcds, part_end = total * (part + 1) / kXcds;
        unsigned long long* counter = b.next_item + part * kQueueStride;
        for (;;) {
            unsigned long long got = 0;
            if (lane == 0) got = atomicAdd(counter, (unsigned long long)kGrab);
            const long long first = part_begin + first_lane(got);
            if (first >= part_end) break;
            const long long last = first + kGrab < part_end ? first + kGrab : part_end;
            // the set of the grab's first item, then step forward across set boundaries
            int k = __builtin_amdgcn_readfirstlane(item_set(b.item_prefix, b.n_sets, first));
            long long begin = b.item_prefix[k], end = b.item_prefix[k + 1];
            for (long long item = first; item < last; item++) {
                while (item >= end) {
                    k++;
                    begin = end;
                    end = b.item_prefix[k + 1];
                }
                pair_item<kBinned>(b, k, item - begin, lane);
            }
        }
    }
}

// Resets the work-queue counters of batch_pairs (what batch_setup also does), so that the pair
// kernel can be launched again without a set-up in between (back-to-back timing).
__global__ void reset_queues(unsigned long long* next_item) {
    if (threadIdx.x < kXcds) next_item[threadIdx.x * kQueueStride] = 0;
}

struct BinWeights {
    float w[kMaxBinNodes];
};

// out[(t * n_c + ic) * n_bins + b] = sum_g w[g] * P(t, ic, b * nodes + g) for one set, whose
// results sit at P[t * plane + ic * row + ie]
__global__ __launch_bounds__(kBlock) void bin_average(const float* __restrict__ P, int n_c, int n_e, int row_stride,
                                                      long long plane, int nodes, BinWeights w,
                                                      float* __restrict__ out) {
    const int n_bins = n_e / nodes;
    const long long total = 9LL * n_c * n_bins;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const long long row = i / n_bins;                 // t * n_c + ic
        const int bin = int(i % n_bins);
        const long long t = row / n_c, ic = row - t * n_c;
        const float* src = P + t * plane + ic * row_stride + (long long)bin * nodes;
        float acc = 0.0f;
        for (int g = 0; g < nodes; g++) acc = __builtin_fmaf(w.w[g], src[g], acc);
        out[i] = acc;
    }
}

}  // namespace

hipError_t launch_bin_average(const float* P, int n_c, int n_e, int row, long long plane, int nodes,
                              const float* weights, float* out, hipStream_t s) {
    if (nodes < 1 || nodes > kMaxBinNodes || n_e % nodes != 0) return hipErrorInvalidValue;
    BinWeights w{};
    for (int g = 0; g < nodes; g++) w.w[g] = weights[g];
    const long long total = 9LL * n_c * (n_e / nodes);
    if (total == 0) return hipSuccess;
    const unsigned blocks = unsigned(std::min<long long>((total + kBlock - 1) / kBlock, 65535));
    hipLaunchKernelGGL(bin_average, dim3(blocks), dim3(kBlock), 0, s, P, n_c, n_e, row, plane, nodes, w, out);
    return hipGetLastError();
}

hipError_t launch_reset_queues(unsigned long long* next_item, hipStream_t s) {
    hipLaunchKernelGGL(reset_queues, dim3(1), dim3(kWave), 0, s, next_item);
    return hipGetLastError();
}

hipError_t launch_batch_pairs(const BatchDesc& b, int blocks, int waves, hipStream_t s) {
    if (blocks <= 0 || blocks % kXcds != 0) return hipErrorInvalidValue;
    const int nodes = b.bin_nodes;
    if (nodes < 0 || nodes > kMaxBinNodes || (nodes & (nodes - 1)) != 0) return hipErrorInvalidValue;
#define CP3_BATCH(W)                                                                                     \
    if (nodes > 0) hipLaunchKernelGGL((batch_pairs<W, true>), dim3(blocks), dim3(W * kWave), 0, s, b);  \
    else hipLaunchKernelGGL((batch_pairs<W, false>), dim3(blocks), dim3(W * kWave), 0, s, b)
    switch (waves) {
        case 4: CP3_BATCH(4); break;
        case 8: CP3_BATCH(8); break;
        case 16: CP3_BATCH(16); break;
        default: return hipErrorInvalidValue;
    }
#undef CP3_BATCH
    return hipGetLastError();
}

}  // namespace cp3
