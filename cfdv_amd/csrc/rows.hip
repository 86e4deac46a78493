// Assembled node rows of K: the element blocks summed per node pair, stored
// as sliced block rows for the solver's products.
//
// The reference multiplies element by element (MPI_GMRES.h:442-566 sums
// R_e x_e over the elements).  A hex8 element holds 64 blocks of 5x5; summed
// per node pair an interior node has 27 blocks, so one product over the
// assembled rows moves 27 * 25 values (and 27 column indices) per node
// instead of 64 * 25 per element: 2.2x fewer bytes for the same operator.
// The rows are formed after every build from the FP64 element blocks (each
// element block is read once) and stored in FP64 and FP32.  Kernels here:
// the assembly, K x (FP64, FP32, FP32 blocks with FP64 vectors), K X for 8
// vectors, and the node-block inverses from the diagonal slots.
//
// Layout.  Rows come in chunks of 64 nodes (lane = node % 64).  Chunk c has
// G_c groups of 4 slots (slot = one neighbour node of the row, the ones with
// most contributing elements first, padded with the row itself and a zero
// block); goff[c] is the
// first group of chunk c.  A group's 4 blocks are 4 * NEQ^2 values per lane,
// row-major per block, stored 16 bytes per lane at a time: value f of group
// g for a lane sits at g * GV * 64 + (f - f % V) * 64 + lane * V + f % V
// (V = 16 / sizeof(T), GV = 4 * NEQ^2), so a wavefront loads 1 KB of
// contiguous memory per 16-byte vector.  cols[(4 g + t) * 64 + lane] is the
// column node of slot t of the group.
//
// Assembly.  For slot k of a chunk, the contributing element blocks of each
// lane are listed in contrib[(coff[k] + i) * 64 + lane] (i < coff[k + 1] -
// coff[k], -1 where a lane has fewer): code = s nnod^2 + N nnod + M for the
// block (N, M) of the element in storage slot s.  The sum runs in list
// order, so the rows are deterministic.
#include "cfdv_internal.h"

namespace cfdv {

namespace {

template <typename T> struct V16;
template <> struct V16<double> { typedef double __attribute__((ext_vector_type(2))) type; };
template <> struct V16<float> { typedef float __attribute__((ext_vector_type(4))) type; };

template <typename T>
__device__ __forceinline__ size_t row_offset(long g, int lane, int f, int GV) {
    constexpr int V = 16 / (int)sizeof(T);
    const int r = f % V;
    return (size_t)g * GV * 64 + (size_t)(f - r) * 64 + (size_t)lane * V + r;
}

// Chunk handled by wave w of block b.  Blocks are dealt to the 8 XCDs
// round-robin; this mapping gives each XCD a contiguous range of chunks so
// that the x entries neighbouring chunks share stay in one L2.
__device__ __forceinline__ long chunk_of_wave(long n_chunks) {
    const long nb = gridDim.x;
    const long b = blockIdx.x;
    const long per = (nb + 7) / 8;
    const long xcd = b % 8, idx = b / 8;
    const long bb = xcd * per + idx;   // < 8 * per
    return bb * (blockDim.x / 64) + threadIdx.x / 64;
}

}  // namespace

// y = K x over the assembled rows; TB blocks, TV vectors, TA accumulation.
// waves_per_eu (2, 2): the registers of two waves per SIMD let the FP64
// instance keep a group's 50 block loads in flight (1.98 -> 1.93 ms at 2.1M
// hex8; the FP32 ones are unchanged)
template <typename TB, typename TV, typename TA, int NEQ>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void rows_matvec_kernel(
    long n_node, long n_chunks, const long* __restrict__ goff, const int* __restrict__ cols,
    const TB* __restrict__ B, const TV* __restrict__ x, TV* __restrict__ y, const int* __restrict__ gate) {
    constexpr int BB = NEQ * NEQ, GV = 4 * BB, V = 16 / (int)sizeof(TB), NVEC = GV / V;
    typedef typename V16<TB>::type vec_t;
    if (gate && *gate) return;
    const long c = chunk_of_wave(n_chunks);
    if (c >= n_chunks) return;
    const int lane = threadIdx.x & 63;
    const long g0 = goff[c], g1 = goff[c + 1];
    TA acc[NEQ];
#pragma unroll
    for (int r = 0; r < NEQ; r++) acc[r] = 0;
    const vec_t* Bv = reinterpret_cast<const vec_t*>(B) + lane;
    for (long g = g0; g < g1; g++) {
        int cj[4];
#pragma unroll
        for (int t = 0; t < 4; t++) cj[t] = cols[(g * 4 + t) * 64 + lane];
        vec_t bv[NVEC];
#pragma unroll
        for (int i = 0; i < NVEC; i++) bv[i] = __builtin_nontemporal_load(Bv + (size_t)(g * NVEC + i) * 64);
        TA xv[4][NEQ];
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int q = 0; q < NEQ; q++) xv[t][q] = (TA)x[(size_t)cj[t] * NEQ + q];
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int r = 0; r < NEQ; r++)
#pragma unroll
                for (int q = 0; q < NEQ; q++) {
                    const int f = t * BB + r * NEQ + q;
                    acc[r] += (TA)bv[f / V][f % V] * xv[t][q];
                }
    }
    const long row = c * 64 + lane;
    if (row < n_node) {
#pragma unroll
        for (int r = 0; r < NEQ; r++) y[(size_t)row * NEQ + r] = (TV)acc[r];
    }
}

// Y = K X for 8 vectors over the FP32 rows: X, Y node-major [n_node * NEQ][8]
// (the 8 values of a dof contiguous, 32 bytes).  One lane per row as in the
// single product; each group's 16-byte block loads serve all 8 vectors, the
// column's 8-wide values come in two 16-byte loads per dof.  The blocks
// differ row by row, so there is no operand two rows share for a
// matrix-core tile; at 4 flop per block byte the product stays HBM-bound on
// the VALU.
template <int NEQ>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 2))) void rows_matvec8_kernel(
    long n_node, long n_chunks, const long* __restrict__ goff, const int* __restrict__ cols,
    const float* __restrict__ B, const float* __restrict__ X, float* __restrict__ Y) {
    constexpr int BB = NEQ * NEQ, GV = 4 * BB, V = 4, NVEC = GV / V;
    typedef typename V16<float>::type vec_t;
    const long c = chunk_of_wave(n_chunks);
    if (c >= n_chunks) return;
    const int lane = threadIdx.x & 63;
    const long g0 = goff[c], g1 = goff[c + 1];
    float acc[NEQ][8];
#pragma unroll
    for (int r = 0; r < NEQ; r++)
#pragma unroll
        for (int v = 0; v < 8; v++) acc[r][v] = 0.0f;
    const vec_t* Bv = reinterpret_cast<const vec_t*>(B) + lane;
    const vec_t* Xv = reinterpret_cast<const vec_t*>(X);
    // per slot only the 16-byte vectors holding its NEQ^2 values are loaded
    // (the vector shared by two slots is read twice, from cache)
    constexpr int NW = (BB + 2 * V - 2) / V;
    for (long g = g0; g < g1; g++) {
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int v0 = (t * BB) / V, off = t * BB - v0 * V;
            vec_t bw[NW];
#pragma unroll
            for (int i = 0; i < NW; i++)
                if (v0 + i < NVEC) bw[i] = __builtin_nontemporal_load(Bv + (size_t)(g * NVEC + v0 + i) * 64);
            const int cj = cols[(g * 4 + t) * 64 + lane];
            vec_t xv[NEQ][2];
#pragma unroll
            for (int q = 0; q < NEQ; q++) {
                xv[q][0] = Xv[((size_t)cj * NEQ + q) * 2];
                xv[q][1] = Xv[((size_t)cj * NEQ + q) * 2 + 1];
            }
#pragma unroll
            for (int r = 0; r < NEQ; r++)
#pragma unroll
                for (int q = 0; q < NEQ; q++) {
                    const int f = off + r * NEQ + q;
                    const float b = bw[f / V][f % V];
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        acc[r][v] += b * xv[q][0][v];
                        acc[r][4 + v] += b * xv[q][1][v];
                    }
                }
        }
    }
    const long row = c * 64 + lane;
    if (row < n_node) {
        vec_t* Yv = reinterpret_cast<vec_t*>(Y);
#pragma unroll
        for (int r = 0; r < NEQ; r++) {
            Yv[((size_t)row * NEQ + r) * 2] = vec_t{acc[r][0], acc[r][1], acc[r][2], acc[r][3]};
            Yv[((size_t)row * NEQ + r) * 2 + 1] = vec_t{acc[r][4], acc[r][5], acc[r][6], acc[r][7]};
        }
    }
}

// rows from the FP64 element blocks.  PAIRS (contexts with rows): each
// node-pair block is NEQ^2 contiguous doubles (pair_base), read whole.
// Otherwise the chunked layout of blk_offset (the entries of an element row
// come in 16-byte pairs).  A block row of NEQ
// entries starting at column j0 = M NEQ is read as whole pairs: NEQ / 2 of
// them for even NEQ, (NEQ + 1) / 2 from the pair holding j0 for odd NEQ
// (one entry of a neighbouring block read and dropped), so each lane's
// loads are full 16-byte vectors.  TR: the rows of K^T -- slot (n, c) sums
// the transposed blocks K_e(c, n), the same contributors with N and M
// swapped.
template <int NNOD, int NEQ, bool TR, bool PAIRS = false>
__global__ __launch_bounds__(256) void rows_assemble_kernel(long n_chunks, const long* __restrict__ goff,
                                                            const long* __restrict__ coff,
                                                            const int* __restrict__ contrib,
                                                            const double* __restrict__ R, double* __restrict__ B64,
                                                            float* __restrict__ B32) {
    constexpr int BB = NEQ * NEQ, GV = 4 * BB, K = NNOD * NEQ, KK = K * K, NN2 = NNOD * NNOD;
    constexpr int NP = (NEQ + 1) / 2;   // pairs read per block row: 3 for NEQ 5, 2 for NEQ 4
    typedef typename V16<double>::type d2_t;
    const d2_t* Rv = reinterpret_cast<const d2_t*>(R);
    const long c = chunk_of_wave(n_chunks);
    if (c >= n_chunks) return;
    const int lane = threadIdx.x & 63;
    const long g0 = goff[c], g1 = goff[c + 1];
    for (long g = g0; g < g1; g++) {
        for (int t = 0; t < 4; t++) {
            const long k = g * 4 + t;
            double acc[BB];
#pragma unroll
            for (int i = 0; i < BB; i++) acc[i] = 0.0;
            // the next contributor's code is in flight with this one's block
            // rows, and all NEQ x NP pair loads issue before the first add
            const long pe = coff[k + 1];
            long p = coff[k];
            int code = p < pe ? contrib[p * 64 + lane] : -1;
            for (; p < pe; p++) {
                const int code_n = p + 1 < pe ? contrib[(p + 1) * 64 + lane] : -1;
                if (PAIRS && code >= 0) {
                    // pair (N, M) of element s: BB contiguous doubles
                    const long s = code / NN2;
                    const int nm = code % NN2;
                    const int N = TR ? nm % NNOD : nm / NNOD, M = TR ? nm / NNOD : nm % NNOD;
                    const double* Rp = R + pair_base<NNOD, NEQ>(s, N * NNOD + M);
                    double w[BB];
#pragma unroll
                    for (int i = 0; i < BB; i++) w[i] = Rp[i];
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int r = 0; r < NEQ; r++)
#pragma unroll
                        for (int q = 0; q < NEQ; q++) acc[TR ? q * NEQ + r : r * NEQ + q] += w[r * NEQ + q];
                } else if (code >= 0) {
                    const long s = code / NN2;
                    const int nm = code % NN2;
                    const int N = TR ? nm % NNOD : nm / NNOD, M = TR ? nm / NNOD : nm % NNOD;
                    const int j0 = M * NEQ, odd = j0 & 1;
                    // pair p of element s sits at (s / 64) 64 KK / 2 + p 64 + s % 64
                    const d2_t* Rs = Rv + (size_t)(s >> 6) * 64 * (KK / 2) + (s & 63);
                    d2_t w[NEQ][NP];
#pragma unroll
                    for (int r = 0; r < NEQ; r++) {
                        const int p0 = ((N * NEQ + r) * K + j0 - odd) >> 1;
#pragma unroll
                        for (int i = 0; i < NP; i++) w[r][i] = Rs[(size_t)(p0 + i) * 64];
                    }
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int r = 0; r < NEQ; r++)
#pragma unroll
                        for (int q = 0; q < NEQ; q++) {
                            const double ve = w[r][q >> 1][q & 1];
                            const int f = TR ? q * NEQ + r : r * NEQ + q;
                            if constexpr (NEQ % 2 == 1) {
                                const double vo = w[r][(q + 1) >> 1][(q + 1) & 1];
                                acc[f] += odd ? vo : ve;
                            } else {   // j0 even: the pairs hold exactly the row
                                acc[f] += ve;
                            }
                        }
                }
                code = code_n;
            }
#pragma unroll
            for (int i = 0; i < BB; i++) {
                const int f = t * BB + i;
                if (B64) B64[row_offset<double>(g, lane, f, GV)] = acc[i];
                if (B32) B32[row_offset<float>(g, lane, f, GV)] = (float)acc[i];
            }
        }
    }
}

// node-block inverses of K's diagonal blocks from the diagonal slot of
// each FP64 row (TR: the rows hold K^T, the slot is read transposed)
template <int NEQ, bool TR>
__global__ __launch_bounds__(256) void nodeblock_rows_kernel(long n_node, const long* __restrict__ diag,
                                                             const double* __restrict__ B64,
                                                             double* __restrict__ P64, float* __restrict__ P32,
                                                             int* __restrict__ status) {
    constexpr int BB = NEQ * NEQ, GV = 4 * BB;
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_node) return;
    const long k = diag[n];
    const long g = k >> 2;
    const int t = (int)(k & 3), lane = (int)(n & 63);
    double A[NEQ][NEQ];
#pragma unroll
    for (int r = 0; r < NEQ; r++)
#pragma unroll
        for (int c = 0; c < NEQ; c++)
            A[r][c] = B64[row_offset<double>(g, lane, t * BB + (TR ? c * NEQ + r : r * NEQ + c), GV)];
    nodeblock_invert_store<NEQ>(A, n, P64, P32, status);
}

void launch_nodeblock_rows(const RowsArgs& a, const long* diag, const double* B64, bool tr, double* P64, float* P32,
                           int* status, hipStream_t st) {
    const int nb = (int)((a.n_node + 255) / 256);
#define L(NQ, T) \
    hipLaunchKernelGGL((nodeblock_rows_kernel<NQ, T>), dim3(nb), dim3(256), 0, st, a.n_node, diag, B64, P64, P32, status)
    if (a.neqn == 5) {
        if (tr) L(5, true);
        else L(5, false);
    } else {
        if (tr) L(4, true);
        else L(4, false);
    }
#undef L
}

// Slot T of group g from the row values val[i] (value f = T NEQ^2 + i of
// the group), stored as whole 16-byte vectors of the row layout (FP64 pairs,
// FP32 quads): a vector whose last values belong to the next slot keeps its
// first ones in c64 / c32 and is written by that slot's call.  The four
// slots of a group are stored in order, and the last vector of a group ends
// with it.  Each wavefront store is one whole kilobyte, stored nontemporal:
// the rows are read again only after the whole assembly, so they need not
// stay in L2 (refined ramp: 22.6 -> 22.0 ms with 1% fewer bytes fetched;
// bit-identical rows).
template <int NEQ, int T>
__device__ __forceinline__ void store_row_slot(long g, int lane, const double* val,
                                               double* __restrict__ B64, float* __restrict__ B32, double& c64,
                                               float (&c32)[3]) {
    constexpr int BB = NEQ * NEQ, GV = 4 * BB, F0 = T * BB, F1 = F0 + BB;
    typedef typename V16<double>::type d2_t;
    typedef typename V16<float>::type f4_t;
    if (B64) {
        d2_t* p = reinterpret_cast<d2_t*>(B64 + (size_t)g * GV * 64) + lane;   // pair q at p + 64 q
#pragma unroll
        for (int q = F0 / 2; 2 * q < F1; q++) {
            if (2 * q + 1 >= F1) {   // completed by the next slot
                c64 = val[2 * q - F0];
                continue;
            }
            const double a = 2 * q < F0 ? c64 : val[2 * q - F0];
            __builtin_nontemporal_store(d2_t{a, val[2 * q + 1 - F0]}, p + (size_t)q * 64);
        }
    }
    if (B32) {
        f4_t* p = reinterpret_cast<f4_t*>(B32 + (size_t)g * GV * 64) + lane;   // quad q at p + 64 q
#pragma unroll
        for (int q = F0 / 4; 4 * q < F1; q++) {
            float w[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int f = 4 * q + u;
                w[u] = f < F0 ? c32[u] : (f < F1 ? (float)val[f - F0] : 0.0f);
            }
            if (4 * q + 3 >= F1) {   // completed by the next slot
#pragma unroll
                for (int u = 0; u < 3; u++) c32[u] = w[u];
                continue;
            }
            __builtin_nontemporal_store(f4_t{w[0], w[1], w[2], w[3]}, p + (size_t)q * 64);
        }
    }
}

// Rows from the node-pair FP64 element blocks, one wavefront (workgroup) per
// chunk of 64 rows.  Contributor i of the 64 row blocks of slot k is read in
// two half rounds of 32 rows, two blocks per load: half-wave h reads the
// block of row 32 r + 2 j + h in half round r (lane = entry, one 200-byte
// run).  The half rounds of all the chunk's slots run as one sequence, and
// the 16 loads of the next half round are issued before the current one is
// summed, so that one is in flight while the wave adds, transposes and
// stores.  The block addresses are formed on the scalar unit from the
// contributor codes (read lane by lane), each lane selecting its half's.
// Each block sums its contributors in list order (the order of
// rows_assemble_kernel, so the rows are the same bit for bit); the sums go
// through LDS to be written in the row layout (lane = row) as whole 16-byte
// vectors.  TR: the rows of K^T, as in rows_assemble_kernel.
template <int NNOD, int NEQ, bool TR>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2))) void rows_assemble_pairs_kernel(
    long n_chunks, const long* __restrict__ goff, const long* __restrict__ coff, const int* __restrict__ contrib,
    const double* __restrict__ R, double* __restrict__ B64, float* __restrict__ B32) {
    constexpr int BB = NEQ * NEQ, NN2 = NNOD * NNOD;
    __shared__ double tile[64][BB];
    const long per = (gridDim.x + 7) / 8;
    const long c = (long)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (c >= n_chunks) return;   // whole workgroup leaves together
    const int lane = threadIdx.x;
    const int half = lane >> 5, sub = lane & 31;
    // entry of the (transposed) block this lane sums
    const int dst = TR ? (sub % NEQ) * NEQ + sub / NEQ : sub;
    const long k0 = goff[c] * 4, k1 = goff[c + 1] * 4;
    const long p0 = coff[k0], p1 = coff[k1];
    // element-block offset of contributor code cd >= 0 (uniform)
    auto block_of = [](unsigned cd) -> size_t {
        const unsigned s = cd / NN2, nm = cd % NN2;
        const unsigned N = TR ? nm % NNOD : nm / NNOD, M = TR ? nm / NNOD : nm % NNOD;
        return pair_base<NNOD, NEQ>((long)s, (int)(N * NNOD + M));
    };
    const bool entry = sub < BB;
    // the 16 loads of half round r with codes `my` (lane l: the code of row
    // l, -1 for none); lanes without a block read R[0] and drop it
    auto issue = [&](int my, int r, double (&v)[16], unsigned& ok) {
        ok = 0;
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int c0 = __builtin_amdgcn_readlane(my, 32 * r + 2 * j);
            const int c1 = __builtin_amdgcn_readlane(my, 32 * r + 2 * j + 1);
            const size_t b0 = block_of((unsigned)max(c0, 0)), b1 = block_of((unsigned)max(c1, 0));
            const bool good = ((half ? c1 : c0) >= 0) & entry;
            ok |= (unsigned)good << j;
            v[j] = R[good ? (half ? b1 : b0) + sub : 0];
        }
    };
    double acc[32];   // acc[16 r + j]: row 32 r + 2 j + half
#pragma unroll
    for (int j = 0; j < 32; j++) acc[j] = 0.0;
    long k = k0;
    double c64 = 0.0;
    float c32[3] = {0.0f, 0.0f, 0.0f};
    // every slot whose contributors end at or before pe: transposed through
    // LDS, stored, the sums cleared
    auto flush_to = [&](long pe) {
        while (k < k1 && coff[k + 1] <= pe) {
            if (sub < BB) {
#pragma unroll
                for (int j = 0; j < 32; j++) tile[32 * (j >> 4) + 2 * (j & 15) + half][dst] = acc[j];
            }
            __syncthreads();
            const double* val = tile[lane];
            const long g = k >> 2;
            switch ((int)(k & 3)) {
                case 0: store_row_slot<NEQ, 0>(g, lane, val, B64, B32, c64, c32); break;
                case 1: store_row_slot<NEQ, 1>(g, lane, val, B64, B32, c64, c32); break;
                case 2: store_row_slot<NEQ, 2>(g, lane, val, B64, B32, c64, c32); break;
                default: store_row_slot<NEQ, 3>(g, lane, val, B64, B32, c64, c32); break;
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < 32; j++) acc[j] = 0.0;
            k++;
        }
    };
    flush_to(p0);   // slots without contributors before the first one
    double va[16], vb[16];
    unsigned oka = 0, okb = 0;
    int code = p0 < p1 ? contrib[p0 * 64 + lane] : -1;
    if (p0 < p1) issue(code, 0, va, oka);
    for (long p = p0; p < p1; p++) {
        const int code_next = p + 1 < p1 ? contrib[(p + 1) * 64 + lane] : -1;
        issue(code, 1, vb, okb);
#pragma unroll
        for (int j = 0; j < 16; j++)
            if ((oka >> j) & 1u) acc[j] += va[j];
        if (p + 1 < p1) issue(code_next, 0, va, oka);
#pragma unroll
        for (int j = 0; j < 16; j++)
            if ((okb >> j) & 1u) acc[16 + j] += vb[j];
        flush_to(p + 1);
        code = code_next;
    }
}

// The rows of K (TR: of K^T) in the product layout from the staged FP64
// rows S [n_stage][NEQ^2] (staged slot (4 g + t) * 64 + lane = slot t of
// group g of the lane's row, one contiguous row-major block each): lane
// `lane` of the wavefront of group g gathers its 4 blocks (for K^T the
// mirrored slots' blocks, transposed; mirror -1: a padding slot, zero) and
// writes the group's 16-byte vectors, each wavefront store one contiguous
// kilobyte.
template <int NEQ, bool TR>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 2))) void rows_from_stage_kernel(
    long n_groups, const double* __restrict__ S, const int* __restrict__ mirror, double* __restrict__ B64,
    float* __restrict__ B32) {
    constexpr int BB = NEQ * NEQ, GV = 4 * BB;
    const long g = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (g >= n_groups) return;
    double v[GV];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const long i = (g * 4 + t) * 64 + lane;
        const long src = TR ? (long)mirror[i] : i;
        const double* b = S + (size_t)(src >= 0 ? src : 0) * BB;
#pragma unroll
        for (int q = 0; q < BB; q++) v[t * BB + q] = src >= 0 ? b[TR ? (q % NEQ) * NEQ + q / NEQ : q] : 0.0;
    }
    typedef double __attribute__((ext_vector_type(2))) d2_t;
    typedef float __attribute__((ext_vector_type(4))) f4_t;
    d2_t* d = reinterpret_cast<d2_t*>(B64 + (size_t)g * GV * 64) + lane;
    f4_t* f = reinterpret_cast<f4_t*>(B32 + (size_t)g * GV * 64) + lane;
    // stored nontemporal, as the assembly's rows (refined ramp: 12.1 ->
    // 11.3 ms); the staged rows are read with ordinary loads, whose lines
    // the wavefront's next loads share
#pragma unroll
    for (int j = 0; j < GV / 2; j++) __builtin_nontemporal_store(d2_t{v[2 * j], v[2 * j + 1]}, d + (size_t)j * 64);
#pragma unroll
    for (int j = 0; j < GV / 4; j++)
        __builtin_nontemporal_store(f4_t{(float)v[4 * j], (float)v[4 * j + 1], (float)v[4 * j + 2], (float)v[4 * j + 3]},
                                    f + (size_t)j * 64);
}

static int row_grid(long n_chunks, int waves_per_block = 4) {
    const long nb = (n_chunks + waves_per_block - 1) / waves_per_block;
    return (int)(((nb + 7) / 8) * 8);   // whole rounds of 8 XCDs
}

void launch_rows_matvec(const RowsArgs& a, const double* B64, const float* B32, const void* x, void* y, int prec,
                        hipStream_t st, const int* gate) {
    if (a.n_chunks <= 0) return;
    // one wavefront per workgroup, so that a chunk with more slots (refined
    // rows) holds no other wave's SIMD slot while it finishes (refined ramp:
    // FP32 2.48 -> 2.46 ms, FP64 4.27 -> 4.16 ms against four per workgroup)
    const dim3 grid(row_grid(a.n_chunks, 1)), blk(64);
#define L(TB, TV, TA, NQ, BP)                                                                              \
    hipLaunchKernelGGL((rows_matvec_kernel<TB, TV, TA, NQ>), grid, blk, 0, st, a.n_node, a.n_chunks, a.goff, \
                       a.cols, BP, (const TV*)x, (TV*)y, gate)
    if (a.neqn == 5) {
        if (prec == 64) L(double, double, double, 5, B64);
        else if (prec == 32) L(float, float, float, 5, B32);
        else L(float, double, double, 5, B32);
    } else {
        if (prec == 64) L(double, double, double, 4, B64);
        else if (prec == 32) L(float, float, float, 4, B32);
        else L(float, double, double, 4, B32);
    }
#undef L
}

void launch_rows_matvec8(const RowsArgs& a, const float* B32, const float* X, float* Y, hipStream_t st) {
    if (a.n_chunks <= 0) return;
    const dim3 grid(row_grid(a.n_chunks)), blk(256);
    if (a.neqn == 5)
        hipLaunchKernelGGL((rows_matvec8_kernel<5>), grid, blk, 0, st, a.n_node, a.n_chunks, a.goff, a.cols, B32, X, Y);
    else
        hipLaunchKernelGGL((rows_matvec8_kernel<4>), grid, blk, 0, st, a.n_node, a.n_chunks, a.goff, a.cols, B32, X, Y);
}

void launch_rows_assemble(const RowsArgs& a, const double* R64, bool tr, double* B64, float* B32, hipStream_t st) {
    if (a.n_chunks <= 0) return;
    const dim3 grid(row_grid(a.n_chunks)), blk(256);
#define L(NN, NQ, T, P)                                                                                        \
    hipLaunchKernelGGL((rows_assemble_kernel<NN, NQ, T, P>), grid, blk, 0, st, a.n_chunks, a.goff, a.coff, a.contrib, \
                       R64, B64, B32)
    if (a.r64_pairs) {
        // one 64-lane workgroup per chunk, whole rounds of the 8 XCDs
        const dim3 pgrid((unsigned)(((a.n_chunks + 7) / 8) * 8)), pblk(64);
#define P(NN, NQ, T)                                                                                               \
    hipLaunchKernelGGL((rows_assemble_pairs_kernel<NN, NQ, T>), pgrid, pblk, 0, st, a.n_chunks, a.goff, a.coff, \
                       a.contrib, R64, B64, B32)
        if (a.nnod == 8) { if (tr) P(8, 5, true); else P(8, 5, false); }
        else { if (tr) P(4, 4, true); else P(4, 4, false); }
#undef P
        return;
    }
    if (a.nnod == 8) {
        if (tr) L(8, 5, true, false); else L(8, 5, false, false);
    } else {
        if (tr) L(4, 4, true, false); else L(4, 4, false, false);
    }
#undef L
}

void launch_rows_from_stage(const RowsArgs& a, long n_stage, const double* S64, const int* mirror, bool tr,
                            double* B64, float* B32, hipStream_t st) {
    if (n_stage <= 0) return;
    const long ng = n_stage / 256;   // groups of 4 slots of 64 rows
    const dim3 grid((unsigned)((ng + 3) / 4)), blk(256);
    if (a.neqn == 5) {
        if (tr) hipLaunchKernelGGL((rows_from_stage_kernel<5, true>), grid, blk, 0, st, ng, S64, mirror, B64, B32);
        else hipLaunchKernelGGL((rows_from_stage_kernel<5, false>), grid, blk, 0, st, ng, S64, mirror, B64, B32);
    } else {
        if (tr) hipLaunchKernelGGL((rows_from_stage_kernel<4, true>), grid, blk, 0, st, ng, S64, mirror, B64, B32);
        else hipLaunchKernelGGL((rows_from_stage_kernel<4, false>), grid, blk, 0, st, ng, S64, mirror, B64, B32);
    }
}

}  // namespace cfdv
