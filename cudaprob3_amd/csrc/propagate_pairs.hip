// Table-driven per-pair propagation.
//
// A layer's amplitude matrix is A = sum_k e^{i phi_k} Q_k with Q_k = U P_k U^dagger the
// flavour-basis projectors of the layer Hamiltonian (the reference's eq. 11 products,
// physics.hpp:346-517, rearranged by linearity).  Q_k depends on (energy, shell) only and
// comes from the spectrum table; per pair and layer only the three phases
// phi_k = -2.534 lambda_k L / E, the sum over k, and the running products
// (physics.hpp:645-677) remain.
//
// FP64 (pairs_f64): A = sum_k (cos phi_k + i sin phi_k) Q_k with the reference's phase
// expression ((-2.534 lambda_k) L) / E, and the reference's running products of full 3x3
// matrices.
//
// Mixed precision (pairs_mixed): changes that are exact in real arithmetic.
//  * A real basis.  The flavour Hamiltonian of every layer is W H_r W^dagger with H_r real
//    symmetric and one unitary W per hypothesis (W = diag(1, W'), it commutes with the
//    electron-only matter term; host_util.hpp real_basis).  So the layer amplitudes are
//    W S W^dagger with S = O D O^T for the layer's real eigenbasis O: complex symmetric, and the
//    mantle product of the way out is the transpose of the product of the way in, so one running
//    product serves both (pair_mixed.hpp).
//  * The layer's global phase is dropped (it cancels in |amplitude|^2): D = diag(1, e^{i dphi_1},
//    e^{i dphi_2}).  The running product is carried in the shells' eigenbases, two of its three
//    columns, and completed from its determinant at the end (pair_mixed.hpp).  dphi_k (up to
//    ~10^3 rad on core-crossing paths at 0.05 GeV, where FP32 would lose 1e-4 rad) is formed and
//    reduced to half a turn plus a sign in FP64; sin/cos of the remainder (polynomials) and the
//    matrix arithmetic are FP32.
//  * The total amplitude is unitary, so its upper-left 2x2 block fixes all nine
//    probabilities (rows and columns of |T|^2 sum to one).
//  * A path that crosses no shell is one vacuum layer: T = U D U^dagger from host constants.

// Mapping: one wavefront = two cosine rows x 64 consecutive energies (lane = energy); the four
// wavefronts of a workgroup take eight consecutive rows of the same energy tile, and
// consecutive logical workgroups stay on one XCD so that an energy tile's table slice is
// reused from that XCD's L2.  The row index is made wave-uniform (readfirstlane), so the
// layer loop runs on scalar control flow.
#include "common.hpp"
#include "pair_mixed.hpp"
#include "reduce.hpp"
#include "sincos_f64.hpp"

namespace cp3 {
namespace {

// A (+)= (c + i s) Q for one Hermitian projector Q packed as (Q00, Q11, Q22, Re/Im Q01, Re/Im Q02,
// Re/Im Q12); first: A = instead of A +=.  Applied for k = 0, 1, 2 in turn, each entry of A is the
// sum over k in that order.
template <typename T>
__device__ inline void accumulate_projector(const T q[9], T c, T s, bool first, Cx<T> A[3][3]) {
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if (first) {
            A[d][d] = Cx<T>{c * q[d], s * q[d]};
        } else {
            A[d][d].re += c * q[d];
            A[d][d].im += s * q[d];
        }
    }
    const int idx[3][2] = {{0, 1}, {0, 2}, {1, 2}};
#pragma unroll
    for (int w = 0; w < 3; w++) {
        const T a = q[3 + 2 * w], b = q[4 + 2 * w];
        const T ca = c * a, sb = s * b, cb = c * b, sa = s * a;
        Cx<T>& u = A[idx[w][0]][idx[w][1]];
        Cx<T>& l = A[idx[w][1]][idx[w][0]];
        if (first) {
            u = Cx<T>{ca - sb, cb + sa};
            l = Cx<T>{ca + sb, sa - cb};
        } else {
            u.re += ca - sb;
            u.im += cb + sa;
            l.re += ca + sb;
            l.im += sa - cb;
        }
    }
}

struct PairArgs {
    const int* maxlayers;
    const double* energies;
    const double* geom;      // [n_c][rows + 1] km
    const void* Q;
    const double* R;
    void* out;               // SoA [9][n_c][n_e]
    int n_c, n_e, rows, ne_pad;
};

// sum_k a[k] b[k] for complex 3-vectors (a row times a column), in the reference's order
// (math.hpp:50-64: the k = 0 term, then k = 1, 2 added).
template <typename T>
__device__ inline Cx<T> cdot3(const Cx<T> a[3], const Cx<T> b[3]) {
    T re = a[0].re * b[0].re - a[0].im * b[0].im;
    T im = a[0].im * b[0].re + a[0].re * b[0].im;
#pragma unroll
    for (int k = 1; k < 3; k++) {
        re += a[k].re * b[k].re - a[k].im * b[k].im;
        im += a[k].im * b[k].re + a[k].re * b[k].im;
    }
    return Cx<T>{re, im};
}

// FP64 pair kernel: R = -2.534 lambda_k, phase = R * L / E.  The way-in product `fin` stays in
// registers and is updated in place a column at a time; the mantle product of every lane waits
// in LDS (park[e][lane], e = (row * 3 + column) * 2 + {re, im}; each lane touches only its own
// column, so no barrier is needed).  With both products and a temporary in registers the kernel
// held 184 registers per lane (two wavefronts per SIMD).
__global__ __launch_bounds__(kBlock) void pairs_f64_kernel(PairArgs a) {
    using T = double;
    __shared__ T park[18][kBlock];
    const unsigned lb = xcd_logical_block(blockIdx.x, gridDim.x);
    const int tiles = (a.n_e + kWave - 1) / kWave;
    const int row_groups = (a.n_c + 3) / 4;
    const int tile = lb / row_groups;
    const int ic = __builtin_amdgcn_readfirstlane((lb % row_groups) * 4 + int(threadIdx.x) / kWave);
    if (tile >= tiles || ic >= a.n_c) return;
    const int ie = tile * kWave + threadIdx.x % kWave;   // < ne_pad: table reads stay in bounds

    const int L = __builtin_amdgcn_readfirstlane(a.maxlayers[ic]);
    const double* gl = a.geom + std::size_t(ic) * (a.rows + 1);
    const T* Q = static_cast<const T*>(a.Q);
    const std::size_t st = a.ne_pad;
    const double E = a.energies[ie < a.n_e ? ie : a.n_e - 1];
    const int lane = int(threadIdx.x);
    auto mantle_get = [&](int x, int y) { return Cx<T>{park[(x * 3 + y) * 2][lane], park[(x * 3 + y) * 2 + 1][lane]}; };
    auto mantle_put = [&](int x, int y, Cx<T> v) {
        park[(x * 3 + y) * 2][lane] = v.re;
        park[(x * 3 + y) * 2 + 1][lane] = v.im;
    };

    Cx<T> fin[3][3], A[3][3];
    for (int i = 0; i <= L; i++) {
        const double dist = gl[i];
        const T* q = Q + std::size_t(i) * 27 * st + ie;
        const double* r = a.R + std::size_t(i) * 3 * st + ie;
        // A = sum_k (c_k + i s_k) Q_k, one projector at a time (the same sums, in the same order,
        // as forming all three phases first; fewer values live at once)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const double arg = r[k * st] * dist / E;
            T ck, sk;
            if (!sincos_reduced_f64(arg, sk, ck)) sincos(arg, &sk, &ck);
            T qk[9];
#pragma unroll
            for (int f = 0; f < 9; f++) qk[f] = q[(k * 9 + f) * st];
            accumulate_projector<T>(qk, ck, sk, k == 0, A);
        }
        if (i == 0) {
#pragma unroll
            for (int x = 0; x < 3; x++)
#pragma unroll
                for (int y = 0; y < 3; y++) fin[x][y] = A[x][y];
            continue;
        }
        // fin <- A fin
#pragma unroll
        for (int y = 0; y < 3; y++) {
            const Cx<T> colv[3] = {fin[0][y], fin[1][y], fin[2][y]};
            Cx<T> out[3];
#pragma unroll
            for (int x = 0; x < 3; x++) out[x] = cdot3<T>(A[x], colv);
#pragma unroll
            for (int x = 0; x < 3; x++) fin[x][y] = out[x];
        }
        if (i >= L) continue;
        // mantle <- mantle A (the first mantle layer is A itself)
#pragma unroll
        for (int x = 0; x < 3; x++) {
            if (i == 1) {
#pragma unroll
                for (int y = 0; y < 3; y++) mantle_put(x, y, A[x][y]);
                continue;
            }
            const Cx<T> rowv[3] = {mantle_get(x, 0), mantle_get(x, 1), mantle_get(x, 2)};
            Cx<T> out[3];
#pragma unroll
            for (int y = 0; y < 3; y++) {
                const Cx<T> colv[3] = {A[0][y], A[1][y], A[2][y]};
                out[y] = cdot3<T>(rowv, colv);
            }
#pragma unroll
            for (int y = 0; y < 3; y++) mantle_put(x, y, out[y]);
        }
    }
    if (L >= 2) {
        // fin <- mantle fin
        Cx<T> mt[3][3];
#pragma unroll
        for (int x = 0; x < 3; x++)
#pragma unroll
            for (int y = 0; y < 3; y++) mt[x][y] = mantle_get(x, y);
#pragma unroll
        for (int y = 0; y < 3; y++) {
            const Cx<T> colv[3] = {fin[0][y], fin[1][y], fin[2][y]};
            Cx<T> out[3];
#pragma unroll
            for (int x = 0; x < 3; x++) out[x] = cdot3<T>(mt[x], colv);
#pragma unroll
            for (int x = 0; x < 3; x++) fin[x][y] = out[x];
        }
    }
    if (ie >= a.n_e) return;
    const std::size_t plane = std::size_t(a.n_c) * a.n_e;
    T* o = static_cast<T*>(a.out) + std::size_t(ic) * a.n_e + ie;
#pragma unroll
    for (int x = 0; x < 3; x++)
#pragma unroll
        for (int y = 0; y < 3; y++) {
            const Cx<T> v = fin[y][x];
            o[(x * 3 + y) * plane] = v.re * v.re + v.im * v.im;
        }
}

// Mixed-precision pair kernel (see the file comment).
struct MixedArgs {
    const int* maxlayers;
    const double* geom;       // [n_c][rows + 1] km
    const float4* Q;          // [(rows + 1) * kMixedQGroups][ne_pad]
    const float4* M;          // [(rows + 1) * kMixedMGroups][ne_pad]
    const double2* R;         // [(rows + 1)][ne_pad]
    float4 wmu;               // row mu of the real basis's W
    VacuumBlock vac;          // U_xk conj(U_yk) for the paths that cross no shell
    float* out;               // SoA [9][n_c][n_e]
    const float* data;        // likelihood variant: SoA [9][n_c][n_e] reference values
    double* partials;         // likelihood variant: one sum per workgroup
    int n_c, n_e, rows, ne_pad;
};

// Each wavefront takes two consecutive cosine rows (kRowsPerWave) at its 64 energies: rows with the
// same number of crossed shells -- all but the few pairs that straddle a shell boundary -- are
// propagated together and share every table read; otherwise one after the other.
//
// kLike: instead of storing P, sum (P - data)^2 in FP64 over the workgroup's pairs (the
// Markov-chain likelihood without a result round trip through HBM); every wavefront takes
// part in the workgroup reduction, so none returns early.
static_assert(kRowsPerWave == 2, "the mixed kernels pair rows two by two");
template <bool kLike>
__global__ __launch_bounds__(kBlock) void pairs_mixed_kernel(MixedArgs a) {
    const unsigned lb = xcd_logical_block(blockIdx.x, gridDim.x);
    const int tiles = (a.n_e + kWave - 1) / kWave;
    constexpr int kRowsPerBlock = kRowsPerWave * (kBlock / kWave);
    const int row_groups = (a.n_c + kRowsPerBlock - 1) / kRowsPerBlock;
    const int tile = lb / row_groups;
    const int ic0 =
        __builtin_amdgcn_readfirstlane(int((lb % row_groups) * kRowsPerBlock + (threadIdx.x / kWave) * kRowsPerWave));
    const int n_rows = tile < tiles ? max(0, min(kRowsPerWave, a.n_c - ic0)) : 0;
    if (!kLike && n_rows == 0) return;
    const int ie = tile * kWave + threadIdx.x % kWave;   // < ne_pad: table reads stay in bounds
    const std::size_t plane = std::size_t(a.n_c) * a.n_e;
    const std::size_t st = a.ne_pad;

    double acc = 0.0;
    auto emit = [&](int ic, const float (&p)[3][3]) {
        if (ie >= a.n_e) return;
        if (kLike) {
            const float* d = a.data + std::size_t(ic) * a.n_e + ie;      // read once: non-temporal
#pragma unroll
            for (int t = 0; t < 9; t++) {
                const double e = double(p[t / 3][t % 3]) - double(__builtin_nontemporal_load(d + t * plane));
                acc = __builtin_fma(e, e, acc);
            }
        } else {
            float* o = a.out + std::size_t(ic) * a.n_e + ie;     // non-temporal: results stream past L2
#pragma unroll
            for (int in = 0; in < 3; in++)
#pragma unroll
                for (int ou = 0; ou < 3; ou++) __builtin_nontemporal_store(p[in][ou], o + (in * 3 + ou) * plane);
        }
    };
    const int L0 = n_rows > 0 ? crossed_shells(a.maxlayers, ic0, a.rows) : 0;
    if (n_rows == 2 && crossed_shells(a.maxlayers, ic0 + 1, a.rows) == L0) {
        const double* gl[2] = {a.geom + std::size_t(ic0) * (a.rows + 1), a.geom + std::size_t(ic0 + 1) * (a.rows + 1)};
        float p[2][3][3];
        propagate_rows_mixed<2, kFormEigenTable>(L0, gl, a.Q, a.M, a.R, st, ie, a.wmu, &a.vac, p);
        emit(ic0, p[0]);
        emit(ic0 + 1, p[1]);
    } else {
#pragma unroll 1
        for (int k = 0; k < n_rows; k++) {
            const int ic = ic0 + k;
            const double* gl[1] = {a.geom + std::size_t(ic) * (a.rows + 1)};
            float p[1][3][3];
            propagate_rows_mixed<1, kFormEigenTable>(crossed_shells(a.maxlayers, ic, a.rows), gl, a.Q, a.M, a.R, st, ie, a.wmu,
                                          &a.vac, p);
            emit(ic, p[0]);
        }
    }
    if (kLike) {
        __shared__ double lds[kBlock / kWave];
        const double s = block_sum(acc, lds);
        if (threadIdx.x == 0) a.partials[blockIdx.x] = s;
    }
}

hipError_t launch_pairs_fp64(const GridDesc& g, int ne_pad, const double* geom, const void* Q, const double* R,
                             void* out, hipStream_t s) {
    if (g.n_c == 0 || g.n_e == 0) return hipSuccess;
    PairArgs a{g.maxlayers, g.energies, geom, Q, R, out, g.n_c, g.n_e, g.rows, ne_pad};
    const unsigned tiles = (g.n_e + kWave - 1) / kWave;
    const unsigned blocks = round_up(tiles * ((g.n_c + 3) / 4), kXcds);
    hipLaunchKernelGGL(pairs_f64_kernel, dim3(blocks), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

}  // namespace

unsigned likelihood_blocks(const GridDesc& g) {
    if (g.n_c == 0 || g.n_e == 0) return 0;
    const unsigned tiles = (g.n_e + kWave - 1) / kWave;
    constexpr int kRowsPerBlock = kRowsPerWave * (kBlock / kWave);
    return round_up(tiles * ((g.n_c + kRowsPerBlock - 1) / kRowsPerBlock), kXcds);
}

hipError_t launch_pairs_mixed(const GridDesc& g, int ne_pad, const double* geom, const float* Q, const float* M,
                              const double* R, float4 wmu, const VacuumBlock& vac, float* out, hipStream_t s) {
    if (g.n_c == 0 || g.n_e == 0) return hipSuccess;
    MixedArgs a{g.maxlayers, geom, reinterpret_cast<const float4*>(Q), reinterpret_cast<const float4*>(M),
                reinterpret_cast<const double2*>(R), wmu, vac, out, nullptr, nullptr, g.n_c, g.n_e, g.rows, ne_pad};
    hipLaunchKernelGGL(pairs_mixed_kernel<false>, dim3(likelihood_blocks(g)), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_pairs_mixed_sqdiff(const GridDesc& g, int ne_pad, const double* geom, const float* Q,
                                     const float* M, const double* R, float4 wmu, const VacuumBlock& vac,
                                     const float* data, double* partials, hipStream_t s) {
    if (g.n_c == 0 || g.n_e == 0) return hipSuccess;
    MixedArgs a{g.maxlayers, geom, reinterpret_cast<const float4*>(Q), reinterpret_cast<const float4*>(M),
                reinterpret_cast<const double2*>(R), wmu, vac, nullptr, data, partials, g.n_c, g.n_e, g.rows, ne_pad};
    hipLaunchKernelGGL(pairs_mixed_kernel<true>, dim3(likelihood_blocks(g)), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_pairs_f64(const GridDesc& g, int ne_pad, const double* geom, const double* Q,
                            const double* R, double* out, hipStream_t s) {
    return launch_pairs_fp64(g, ne_pad, geom, Q, R, out, s);
}

}  // namespace cp3
