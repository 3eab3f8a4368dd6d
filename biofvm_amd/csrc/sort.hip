// Device sort and scan for the agent grouping (agents.hip): a stable
// least-significant-digit radix sort of (key, value) pairs, 8 bits per pass,
// and an exclusive prefix sum of ints.  Integer work only, so the results do
// not depend on the launch shapes; the sort keeps equal keys in input order,
// which is what keeps an agent list's order inside a voxel
// (BioFVM_basic_agent.cpp:229-238 applies a voxel's agents in list order).
//
// Sort pass (digit d of bits [shift, shift+8) of the key):
//   k_digit_hist     per block of kSortTile pairs, the count of each digit,
//                    written digit-major: hist[d * nblocks + block];
//   exclusive_scan   of hist: the first output slot of (digit, block), in
//                    digit order and, within a digit, block order;
//   k_digit_scatter  each block walks its tile in order, kSortThreads pairs
//                    at a time, and places a pair after every earlier pair of
//                    its digit: earlier chunks (a running count per digit),
//                    earlier waves of the chunk (per-wave counts), earlier
//                    lanes of its wave (lanes with the same digit, found with
//                    one ballot per digit bit).
// Scan: per block of kScanTile values an exclusive scan and the block's total
// (k_scan_tiles), the totals scanned by one block in chunks with a carry
// (k_scan_totals), the totals added to the blocks' values (k_scan_add).
#include <algorithm>

#include "bfvm_internal.h"

namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kDigits = 256;
constexpr int kWaves = kSortThreads / 64;

constexpr int kScanThreads = 1024;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

__global__ void __launch_bounds__(kSortThreads) k_digit_hist(const unsigned* __restrict__ key, int n, int shift,
                                                             int nblocks, int* __restrict__ hist)
{
    __shared__ int cnt[kDigits];
    for (int d = threadIdx.x; d < kDigits; d += kSortThreads) cnt[d] = 0;
    __syncthreads();
    const int first = blockIdx.x * kSortTile;
    for (int q = threadIdx.x; q < kSortTile; q += kSortThreads) {
        const int i = first + q;
        if (i < n) atomicAdd(&cnt[(key[i] >> shift) & (kDigits - 1)], 1);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < kDigits; d += kSortThreads) hist[d * nblocks + blockIdx.x] = cnt[d];
}

__global__ void __launch_bounds__(kSortThreads)
k_digit_scatter(const unsigned* __restrict__ key, const int* __restrict__ val, int n, int shift, int nblocks,
                const int* __restrict__ offset, unsigned* __restrict__ key_out, int* __restrict__ val_out)
{
    __shared__ int run[kDigits];             // pairs of each digit in this block's earlier chunks
    __shared__ int wave_cnt[kWaves][kDigits];  // pairs of each digit in each wave of this chunk
    const int tid = threadIdx.x, wave = tid / 64;
    const unsigned long long below = __lanemask_lt();
    for (int d = tid; d < kDigits; d += kSortThreads) run[d] = 0;
    const int first = blockIdx.x * kSortTile;
    for (int c = 0; c < kSortItems; ++c) {
        for (int d = tid; d < kDigits; d += kSortThreads)
            for (int w = 0; w < kWaves; ++w) wave_cnt[w][d] = 0;
        __syncthreads();
        const int i = first + c * kSortThreads + tid;
        const bool valid = i < n;
        const unsigned k = valid ? key[i] : 0u;
        const int d = valid ? (int)((k >> shift) & (kDigits - 1)) : kDigits;  // kDigits: no pair
        // lanes of this wave with the same digit (9 bits: the invalid lanes
        // form a group of their own)
        unsigned long long peers = ~0ull;
        for (int b = 0; b < 9; ++b) {
            const bool bit = (d >> b) & 1;
            const unsigned long long set = __ballot(bit);
            peers &= bit ? set : ~set;
        }
        const int before_in_wave = __popcll(peers & below);
        if (valid && before_in_wave == 0) wave_cnt[wave][d] = __popcll(peers);
        __syncthreads();
        if (valid) {
            int rank = run[d] + before_in_wave;
            for (int w = 0; w < wave; ++w) rank += wave_cnt[w][d];
            const int dst = offset[d * nblocks + blockIdx.x] + rank;
            key_out[dst] = k;
            val_out[dst] = val[i];
        }
        __syncthreads();
        for (int dd = tid; dd < kDigits; dd += kSortThreads) {
            int add = 0;
            for (int w = 0; w < kWaves; ++w) add += wave_cnt[w][dd];
            run[dd] += add;
        }
        // the next chunk clears wave_cnt after this barrier
        __syncthreads();
    }
}

// Exclusive scan of a block's kScanThreads partial sums in LDS (part), in
// place; returns the block total.
__device__ int block_exclusive_scan(int* part)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    __shared__ int wave_sum[kScanThreads / 64];
    int v = part[tid];
    // inclusive scan within the wave
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    if (lane == 63) wave_sum[wave] = v;
    __syncthreads();
    int before = 0;
    for (int w = 0; w < wave; ++w) before += wave_sum[w];
    int total = 0;
    for (int w = 0; w < kScanThreads / 64; ++w) total += wave_sum[w];
    const int exclusive = before + v - part[tid];
    __syncthreads();
    part[tid] = exclusive;
    __syncthreads();
    return total;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_tiles(const int* __restrict__ in, int* __restrict__ out, int n,
                                                             int* __restrict__ totals)
{
    __shared__ int part[kScanThreads];
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems], s = 0;
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) {
        v[q] = base + q < n ? in[base + q] : 0;
        s += v[q];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    const int total = block_exclusive_scan(part);
    int run = part[threadIdx.x];
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) {
        if (base + q < n) out[base + q] = run;
        run += v[q];
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = total;
}

// One block: exclusive scan of the block totals, kScanThreads at a time.
__global__ void __launch_bounds__(kScanThreads) k_scan_totals(int* __restrict__ totals, int nb)
{
    __shared__ int part[kScanThreads];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int c = 0; c < nb; c += kScanThreads) {
        const int i = c + threadIdx.x;
        part[threadIdx.x] = i < nb ? totals[i] : 0;
        __syncthreads();
        const int total = block_exclusive_scan(part);
        if (i < nb) totals[i] = carry + part[threadIdx.x];
        __syncthreads();
        if (threadIdx.x == 0) carry += total;
        __syncthreads();
    }
}

__global__ void k_scan_add(int* __restrict__ out, int n, const int* __restrict__ totals)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += totals[i / kScanTile];
}

int scan_blocks(int n) { return (n + kScanTile - 1) / kScanTile; }
int sort_blocks(int n) { return (n + kSortTile - 1) / kSortTile; }

}  // namespace

namespace bfvm {

size_t exclusive_scan_tmp_bytes(int n)
{
    return sizeof(int) * (size_t)std::max(1, scan_blocks(n));
}

hipError_t exclusive_scan(const int* in, int* out, int n, void* tmp, size_t tmp_bytes, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    const int nb = scan_blocks(n);
    if (tmp_bytes < exclusive_scan_tmp_bytes(n)) return hipErrorInvalidValue;
    int* totals = (int*)tmp;
    k_scan_tiles<<<nb, kScanThreads, 0, st>>>(in, out, n, totals);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_scan_totals<<<1, kScanThreads, 0, st>>>(totals, nb);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_scan_add<<<(n + 255) / 256, 256, 0, st>>>(out, n, totals);
    return hipGetLastError();
}

size_t radix_sort_tmp_bytes(int n)
{
    const int nb = std::max(1, sort_blocks(n));
    const size_t hist = sizeof(int) * (size_t)kDigits * nb;
    // two key and value buffers, the digit counts, their scan and its workspace
    return 2 * (sizeof(unsigned) + sizeof(int)) * (size_t)std::max(n, 1) + 2 * hist +
           exclusive_scan_tmp_bytes(kDigits * nb) + 256;
}

hipError_t radix_sort_pairs(const unsigned* key_in, const int* val_in, unsigned* key_out, int* val_out, int n,
                            int bits, void* tmp, size_t tmp_bytes, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    if (tmp_bytes < radix_sort_tmp_bytes(n) || bits < 1 || bits > 32) return hipErrorInvalidValue;
    const int nb = sort_blocks(n);
    char* p = (char*)tmp;
    auto carve = [&](size_t bytes) {
        char* q = p;
        p += (bytes + 255) / 256 * 256;
        return q;
    };
    unsigned* ka = (unsigned*)carve(sizeof(unsigned) * n);
    int* va = (int*)carve(sizeof(int) * n);
    unsigned* kb = (unsigned*)carve(sizeof(unsigned) * n);
    int* vb = (int*)carve(sizeof(int) * n);
    int* hist = (int*)carve(sizeof(int) * (size_t)kDigits * nb);
    int* offs = (int*)carve(sizeof(int) * (size_t)kDigits * nb);
    int* scan_tmp = (int*)carve(exclusive_scan_tmp_bytes(kDigits * nb));
    const int passes = (bits + 7) / 8;
    const unsigned* ks = key_in;
    const int* vs = val_in;
    for (int pass = 0; pass < passes; ++pass) {
        // the last pass writes the caller's buffers, the others alternate
        unsigned* kd = pass == passes - 1 ? key_out : (pass % 2 == 0 ? ka : kb);
        int* vd = pass == passes - 1 ? val_out : (pass % 2 == 0 ? va : vb);
        const int shift = 8 * pass;
        k_digit_hist<<<nb, kSortThreads, 0, st>>>(ks, n, shift, nb, hist);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        if ((e = exclusive_scan(hist, offs, kDigits * nb, scan_tmp, exclusive_scan_tmp_bytes(kDigits * nb), st)))
            return e;
        k_digit_scatter<<<nb, kSortThreads, 0, st>>>(ks, vs, n, shift, nb, offs, kd, vd);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        ks = kd;
        vs = vd;
    }
    return hipSuccess;
}

}  // namespace bfvm
