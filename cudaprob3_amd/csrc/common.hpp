// Shared definitions of the MI355X propagation library (device and host).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace cp3 {

// Physical constants of the reference (constants.hpp:28-40, physics.hpp:442).
constexpr double kGF = 1.52588e-4;          // 2*sqrt(2)*G_F*N_A scale, eV^2 per (GeV g/cm^3)
constexpr double kKm2Cm = 1.0e5;
constexpr double kREarth = 6371.0;          // km
constexpr double kREarthCm = kREarth * kKm2Cm;
constexpr double kYe = 0.5;                 // electrons per nucleon
constexpr double kLoE = 2.534;              // (1/2)(1/(hbar c)) in GeV/(eV^2 km)
constexpr double kPi = 3.14159265358979323846;  // M_PI

// Everything a launch needs about the oscillation hypothesis, passed by value as a
// kernel argument (no process-global device symbols, so concurrent propagators with
// different parameters never race).  U is row-major, (re, im) interleaved.
struct OscParams {
    double U[18];      // PMNS matrix U(i,j) = (U[2(3i+j)], U[2(3i+j)+1])
    double DM[9];      // DM(i,j) = m_i^2 - m_j^2 (with the reference's degeneracy nudge)
    int order[3];      // matter root k that continues vacuum mass i
    int antinu;        // 1: matter potential sign flipped (the only nu/nubar difference)
    // real basis of the mixed kernels (host_util.hpp real_basis): the flavour Hamiltonian is
    // W (Hr - fac e_0 e_0^T) W^dagger with Hr real symmetric and W = diag(1, W')
    double Hr[6];      // (H00, H11, H22, H01, H02, H12)
    double Wmu[4];     // row 1 of W: (Re W11, Im W11, Re W12, Im W12)
};

constexpr int kWave = 64;            // CDNA wavefront
constexpr int kBlock = 256;          // 4 wavefronts per workgroup
constexpr int kRowsPerWave = 2;      // cosine rows per wavefront (single-grid) or work item (batched), mixed kernels
constexpr int kXcds = 8;             // MI355X accelerator complex dies

// Workgroup id -> logical id so that consecutive logical ids run on the same XCD
// (hardware hands workgroups to XCDs round robin).  grid must be a multiple of kXcds.
__device__ inline unsigned xcd_logical_block(unsigned b, unsigned nblocks) {
    const unsigned per = nblocks / kXcds;
    return (b % kXcds) * per + (b / kXcds);
}

inline unsigned round_up(unsigned x, unsigned m) { return (x + m - 1) / m * m; }

// Shells crossed by cosine row ic, kept in [0, rows] so that no table, layer-length or radius
// read leaves its array whatever the caller passed (entry points that validate maxlayers on the
// host never trigger the clamp; cp3h_propagate_mixed_noalloc takes maxlayers in device memory
// and cannot validate them).
__device__ inline int crossed_shells(const int* maxlayers, int ic, int rows) {
    return min(max(maxlayers[ic], 0), rows);
}

// ---- launchers (defined in the .hip translation units) --------------------------------

struct GridDesc {
    const double* cosines;     // [n_c]
    const double* energies;    // [n_e]
    const int* maxlayers;      // [n_c], each in [0, rows]
    const double* radii;       // [rows] descending, radii[0] = outer radius
    const double* rhos;        // [rows] density of the shell whose outer radius is radii[i]
    int n_c, n_e, rows;
    double h_cm;               // production height in cm
};

// Plain FP64 kernel: the reference's per-pair Barger step in double.  out: SoA [9][n_c][n_e].
hipError_t launch_plain_f64(const GridDesc& g, const OscParams& p, double* out, hipStream_t s);

// Layer slots: layer i of a path (0 = atmosphere, i >= 1 = the i-th shell from the surface)
// uses density slot i; maxlayers <= rows (== rows only for a model without an r = 0 row),
// so rows + 1 slots cover every path.

// Per-cosine layer lengths in km: geom[ic*(rows+1) + i], i = 0..maxlayers[ic].
hipError_t launch_geometry(const GridDesc& g, double* geom, hipStream_t s);

// Per-(shell, energy) spectrum tables, s = 0..rows (s = 0 is vacuum, s >= 1 has the density
// rhos[s-1]); ne_pad = n_e rounded up to 64.
//  FP64:  Q[(s*27 + k*9 + f)*ne_pad + e] (Q_k packed, 9 reals), R[(s*3 + k)*ne_pad + e] =
//         -2.534 lambda_k (phase R*L/E).
//  mixed: Qm float4 groups [(s*2 + g)*ne_pad + e] holding the real unit eigenvectors v_1, v_2
//         of the layer Hamiltonian in the real basis (v_1, then v_2, then 2 zero),
//         Mm float4 groups [(s*3 + r)*ne_pad + e] holding row r of the basis change
//         M_s = O_s^T O_{s-1} between the eigenbases O = [v_1 x v_2, v_1, v_2] of slots s and
//         s-1 (O_0 taken as I; M_0 unused), and
//         Rm[(s*ne_pad + e)*2 + k-1] = -2.534 (lambda_k - lambda_0) / E in half turns per km.
constexpr int kMixedQGroups = 2;
constexpr int kMixedQFloats = 4 * kMixedQGroups;
constexpr int kMixedMGroups = 3;
constexpr int kMixedMFloats = 4 * kMixedMGroups;
constexpr int kMixedRDoubles = 2;
// bytes of the mixed tables per (slot, energy) cell
constexpr int kMixedCellBytes = kMixedQFloats * 4 + kMixedMFloats * 4 + kMixedRDoubles * 8;
// Mixed set-up in one launch: layer lengths (geom) and the mixed spectrum tables.
hipError_t launch_setup_mixed(const GridDesc& g, const OscParams& p, int ne_pad, double* geom, float* Qm, float* Mm,
                              double* Rm, hipStream_t s);
hipError_t launch_spectrum_f64(const GridDesc& g, const OscParams& p, int ne_pad, double* Q, double* R, hipStream_t s);

// Vacuum amplitude constants of the single-grid mixed kernel (a path crossing no shell is one
// vacuum layer, whose amplitude is delta_xy + sum_{k=1,2} (e^{i dphi_k} - 1) U_xk conj(U_yk)):
// c[(k - 1) * 8 + (x * 2 + y) * 2 + {0: re, 1: im}] = U_xk conj(U_yk), x, y in (e, mu), float.
struct VacuumBlock {
    float c[16];
};

// Table-driven per-pair kernels.
// wmu: row 1 of the real basis's W in float (OscParams::Wmu); vac: the vacuum constants
hipError_t launch_pairs_mixed(const GridDesc& g, int ne_pad, const double* geom, const float* Q, const float* M,
                              const double* R, float4 wmu, const VacuumBlock& vac, float* out, hipStream_t s);
hipError_t launch_pairs_f64(const GridDesc& g, int ne_pad, const double* geom, const double* Q,
                            const double* R, double* out, hipStream_t s);

// ---- device reductions --------------------------------------------------------------------

// Number of per-workgroup partial sums launch_sqdiff_sum needs as scratch (plus one slot).
int reduce_partials_needed();
// sum of n device doubles in FP64 (fixed order); scratch: reduce_partials_needed() + 1
// doubles, the sum lands in scratch[reduce_partials_needed()].
hipError_t launch_sum_f64(const double* in, long long n, double* scratch, hipStream_t s);
// Likelihood variant of the mixed pair kernel: per workgroup, the FP64 sum of (P - data)^2 over
// its pairs and nine channels, into partials[likelihood_blocks(g)]; nothing else is stored.
unsigned likelihood_blocks(const GridDesc& g);
hipError_t launch_pairs_mixed_sqdiff(const GridDesc& g, int ne_pad, const double* geom, const float* Q,
                                     const float* M, const double* R, float4 wmu, const VacuumBlock& vac,
                                     const float* data, double* partials, hipStream_t s);
// sum_i (a_i - b_i)^2 in FP64 for device floats; result in partials[reduce_partials_needed()].
hipError_t launch_sqdiff_sum(const float* a, const float* b, std::size_t n, double* partials, hipStream_t s);

// ---- batched parameter scans ------------------------------------------------------------

// One parameter set of a batch: its hypothesis and where its grid, Earth model, layer
// lengths, spectrum table and results sit in the batch's concatenated device arrays.
struct BatchSet {
    OscParams p;
    double h_cm;
    long long cos_off;    // cosines / maxlayers
    long long e_off;      // energies
    long long model_off;  // radii / rhos (descending)
    long long geom_off;   // layer lengths, n_c * (rows + 1)
    long long tab_off;    // spectrum cells, (rows + 1) * ne_pad
    long long out_off;    // results: P(t, ic, ie) at out[out_off + t * out_plane + ic * out_row + ie]
    long long out_plane;  // n_c * out_row
    int out_row;          // n_e rounded up to kOutRowAlign floats, so that every row starts a cache line
    // binned mode: bin average (t, ic, bin) -> bin_out[bin_off + t * bin_plane + ic * bin_row + bin]
    long long bin_off, bin_plane;
    int bin_row;
    int has_limits;       // row ic computes only energies below row_limit[cos_off + ic]
    int n_c, n_e, rows, ne_pad;
};

constexpr int kMaxBinNodes = 32;
constexpr int kOutRowAlign = 32;     // floats: 128 bytes, one cache line
// work-queue counters of the batched pair kernel, one per XCD, each in its own 256-byte line
constexpr int kQueueStride = 32;

struct BatchDesc {
    const BatchSet* sets;
    int n_sets;
    const long long* geom_prefix;   // [n_sets + 1] prefix sums of n_c * (rows + 1)
    const long long* spec_prefix;   // [n_sets + 1] prefix sums of (rows + 1) * ne_pad
    const long long* item_prefix;   // [n_sets + 1] prefix sums of n_c * ceil(n_e / 64)
    const double* cosines;
    const int* maxlayers;
    const double* energies;
    const double* radii;
    const double* rhos;
    double* geom;
    float4* Q;                      // kMixedQGroups float4 per spectrum cell (no basis changes: the pair kernel forms them)
    double2* R;
    unsigned long long* next_item;  // kXcds work-queue counters of the pair kernel, kQueueStride apart
                                    // (zeroed by the set-up kernel)
    const int* row_limit;           // per cosine row (parallel to cosines): energies to compute
    float* out;                     // unbinned results (bin_nodes == 0)
    float* bin_out;                 // binned mode: destination of the bin averages
    int bin_nodes;                  // 0, or quadrature nodes per energy bin (a power of two <= 32)
    float bin_w[kMaxBinNodes];      // quadrature weights of the nodes (summing to 1)
};

__device__ inline GridDesc batch_grid(const BatchDesc& b, const BatchSet& s) {
    return GridDesc{b.cosines + s.cos_off, b.energies + s.e_off, b.maxlayers + s.cos_off,
                    b.radii + s.model_off, b.rhos + s.model_off, s.n_c, s.n_e, s.rows, s.h_cm};
}

// Set-up of every set (layer lengths + mixed spectrum tables), grid-stride over all cells.
hipError_t launch_batch_setup(const BatchDesc& b, int blocks, hipStream_t s);
// Persistent pair kernel over all sets: `blocks` workgroups of `waves` wavefronts
// (waves in {4, 8, 16}); each wavefront takes (set, cosine row, 64-energy tile) items.  With
// b.bin_nodes > 0 each set's energies are bin_nodes quadrature nodes per bin and only the
// weighted bin averages are stored.
hipError_t launch_batch_pairs(const BatchDesc& b, int blocks, int waves, hipStream_t s);
// Zeroes the pair kernel's work-queue counters (kXcds of them, kQueueStride apart).
hipError_t launch_reset_queues(unsigned long long* next_item, hipStream_t s);

// Energy-bin average of one set's results (P(t, ic, ie) at P[t * plane + ic * row + ie]) whose
// n_e energies are `nodes` quadrature nodes per bin: out[9][n_c][n_e / nodes].
hipError_t launch_bin_average(const float* P, int n_c, int n_e, int row, long long plane, int nodes,
                              const float* weights, float* out,
                              hipStream_t s);

}  // namespace cp3
