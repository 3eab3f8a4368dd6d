// Reductions of the density store: per-substrate sums in FP64 (the mass
// balance of the LOD step) and the count of values that are not finite.
//
// Every column of each sweep matrix sums to 1 + c2 (the reference's
// tridiagonal matrix, BioFVM_solvers.cpp:121-135: diagonal 1 + 2 c1 + c2 with
// end entries 1 + c1 + c2, off-diagonals -c1), so without sources, Dirichlet
// nodes or agents one LOD_3D step multiplies the sum of each substrate by
// (1 + c2)^-3, c2 = dt*lambda/3.  The sums are taken in a fixed order (block
// b owns a fixed voxel range, its threads a fixed stride, partials folded by
// a fixed tree), so the same store gives the same bits on every run.
#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "bfvm_internal.h"

using bfvm::Env;

namespace bfvm {
namespace {

constexpr int kRedBlock = 256;
constexpr int kMaxRedBlocks = 1024;

// Blocks of the partial pass for nvox voxels: a fixed function of nvox.
int red_blocks(size_t nvox)
{
    return (int)std::max<size_t>(1, std::min<size_t>(kMaxRedBlocks, (nvox + 255) / 256));
}

// Block b sums voxels [b*vpb, (b+1)*vpb) per substrate into part[b*S + s]
// and counts non-finite values into bad[b].  Substrates are taken kRedBlock
// at a time: with sw of them, thread t < (kRedBlock/sw)*sw keeps substrate
// t % sw of every (kRedBlock/sw)-th voxel, so consecutive threads read
// consecutive values.
template <typename T>
__global__ void __launch_bounds__(kRedBlock) k_mass_partials(const T* __restrict__ rho, size_t nvox, int S,
                                                             size_t vpb, double* __restrict__ part,
                                                             unsigned long long* __restrict__ bad)
{
    __shared__ double acc[kRedBlock];
    __shared__ unsigned long long nb[kRedBlock];
    const int tid = threadIdx.x;
    const size_t v0 = (size_t)blockIdx.x * vpb;
    const size_t v1 = v0 + vpb < nvox ? v0 + vpb : nvox;
    unsigned long long cnt = 0;
    for (int s0 = 0; s0 < S; s0 += kRedBlock) {
        const int sw = min(kRedBlock, S - s0);
        const int per = kRedBlock / sw;  // voxels per round
        double sum = 0.0;
        if (tid < per * sw) {
            const int s = s0 + tid % sw;
            for (size_t v = v0 + (size_t)(tid / sw); v < v1; v += (size_t)per) {
                const double x = (double)rho[v * (size_t)S + (size_t)s];
                sum += x;
                cnt += __builtin_isfinite(x) ? 0ull : 1ull;
            }
        }
        acc[tid] = sum;
        __syncthreads();
        if (tid < sw) {
            double t = 0.0;
            for (int k = 0; k < per; ++k) t += acc[k * sw + tid];
            part[(size_t)blockIdx.x * S + s0 + tid] = t;
        }
        __syncthreads();
    }
    nb[tid] = cnt;
    __syncthreads();
    for (int w = kRedBlock / 2; w > 0; w >>= 1) {
        if (tid < w) nb[tid] += nb[tid + w];
        __syncthreads();
    }
    if (tid == 0) bad[blockIdx.x] = nb[0];
}

// One block: sums[s] = sum over the nblk partials of substrate s, and
// nonfinite = sum of the per-block counts, each by the same fixed tree.
__global__ void __launch_bounds__(kRedBlock) k_mass_final(const double* __restrict__ part,
                                                          const unsigned long long* __restrict__ bad,
                                                          int nblk, int S, double* __restrict__ sums,
                                                          long long* __restrict__ nonfinite)
{
    __shared__ double acc[kRedBlock];
    __shared__ unsigned long long nb[kRedBlock];
    const int tid = threadIdx.x;
    for (int s = 0; s < S; ++s) {
        double t = 0.0;
        for (int b = tid; b < nblk; b += kRedBlock) t += part[(size_t)b * S + s];
        acc[tid] = t;
        __syncthreads();
        for (int w = kRedBlock / 2; w > 0; w >>= 1) {
            if (tid < w) acc[tid] += acc[tid + w];
            __syncthreads();
        }
        if (tid == 0) sums[s] = acc[0];
        __syncthreads();
    }
    unsigned long long c = 0;
    for (int b = tid; b < nblk; b += kRedBlock) c += bad[b];
    nb[tid] = c;
    __syncthreads();
    for (int w = kRedBlock / 2; w > 0; w >>= 1) {
        if (tid < w) nb[tid] += nb[tid + w];
        __syncthreads();
    }
    if (tid == 0) *nonfinite = (long long)nb[0];
}

}  // namespace

// Workspace of one mass balance: block partials (nblk*S doubles), per-block
// counts (nblk), then the S sums and the count handed back.
MassLayout mass_layout(const Env& e, void* workspace)
{
    MassLayout m;
    m.nblk = red_blocks(e.nvox);
    m.part = (double*)workspace;
    m.bad = (unsigned long long*)(m.part + (size_t)m.nblk * e.S);
    m.sums = (double*)(m.bad + m.nblk);
    m.count = (long long*)(m.sums + e.S);
    m.bytes = sizeof(double) * ((size_t)m.nblk * e.S + e.S) + sizeof(unsigned long long) * (m.nblk + 1);
    return m;
}

hipError_t launch_mass_balance(Env& e, const MassLayout& m, double* sums, long long* nonfinite)
{
    const size_t vpb = (e.nvox + m.nblk - 1) / m.nblk;
    if (e.precision == BFVM_FP64)
        k_mass_partials<double><<<m.nblk, kRedBlock, 0, e.stream>>>((const double*)e.rho, e.nvox, e.S, vpb,
                                                                   m.part, m.bad);
    else
        k_mass_partials<float><<<m.nblk, kRedBlock, 0, e.stream>>>((const float*)e.rho, e.nvox, e.S, vpb,
                                                                  m.part, m.bad);
    hipError_t st = hipGetLastError();
    if (st != hipSuccess) return st;
    k_mass_final<<<1, kRedBlock, 0, e.stream>>>(m.part, m.bad, m.nblk, e.S, sums, nonfinite);
    return hipGetLastError();
}

}  // namespace bfvm

namespace {
int rfail(Env* e, int code, const std::string& msg)
{
    if (e) e->err = msg;
    return code;
}
int rhip(Env* e, hipError_t st, const char* what)
{
    return rfail(e, BFVM_E_DEVICE, std::string(what) + ": " + bfvm::hip_error_text(st));
}
}  // namespace

extern "C" int bfvm_mass_balance(bfvm_handle h, double* sums, long long* nonfinite)
{
    bfvm::RandGuard rand_guard;
    if (!h) return rfail(nullptr, BFVM_E_ARG, "null handle");
    if (!sums && !nonfinite) return rfail(h, BFVM_E_ARG, "sums and nonfinite are both NULL");
    Env& e = *h;
    hipError_t st;
    if (e.ensemble_stream && (st = hipStreamSynchronize(e.ensemble_stream)) != hipSuccess)
        return rhip(h, st, "bfvm_mass_balance");
    bfvm::MassLayout m = bfvm::mass_layout(e, nullptr);
    if (e.red_bytes < m.bytes) {
        void* p = nullptr;
        if ((st = hipMalloc(&p, m.bytes)) != hipSuccess) return rhip(h, st, "bfvm_mass_balance");
        if (e.red) (void)hipFree(e.red);
        e.red = p;
        e.red_bytes = m.bytes;
    }
    m = bfvm::mass_layout(e, e.red);
    std::vector<double> hs(e.S);
    long long hb = 0;
    if ((st = bfvm::launch_mass_balance(e, m, m.sums, m.count)) != hipSuccess ||
        (st = hipStreamSynchronize(e.stream)) != hipSuccess ||
        (st = hipMemcpy(hs.data(), m.sums, sizeof(double) * e.S, hipMemcpyDeviceToHost)) != hipSuccess ||
        (st = hipMemcpy(&hb, m.count, sizeof(long long), hipMemcpyDeviceToHost)) != hipSuccess)
        return rhip(h, st, "bfvm_mass_balance");
    if (sums) std::copy(hs.begin(), hs.end(), sums);
    if (nonfinite) *nonfinite = hb;
    return BFVM_OK;
}

// ---------------------------------------------------------------------------
// bfvm_mass_balance_noalloc: the same reduction with no allocation.  Its
// workspace is `scratch`, a file-scope device array (one copy per GPU, held
// until the process ends), sized for up to kNoAllocMaxS substrates with the
// largest block count.  Host code never passes `scratch` itself (in host
// code its name stands for a host-side shadow, not a device address): the
// device address is looked up once per GPU, with that GPU current, and kept.
namespace {

constexpr int kNoAllocMaxS = 64;
constexpr size_t kScratchBytes = sizeof(double) * ((size_t)bfvm::kMaxRedBlocks * kNoAllocMaxS + kNoAllocMaxS) +
                                 sizeof(unsigned long long) * (bfvm::kMaxRedBlocks + 1);
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device` (looked up on the first call for
// that device).
hipError_t scratch_address(int device, void** out)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = 0;
        hipError_t st = hipGetDevice(&cur);
        if (st != hipSuccess) return st;
        if (cur != device && (st = hipSetDevice(device)) != hipSuccess) return st;
        void* p = nullptr;
        st = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) {
            const hipError_t st2 = hipSetDevice(cur);
            if (st == hipSuccess) st = st2;
        }
        if (st != hipSuccess) return st;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

}  // namespace

extern "C" int bfvm_mass_balance_noalloc(bfvm_handle h, double* sums, long long* nonfinite)
{
    bfvm::RandGuard rand_guard;
    if (!h) return rfail(nullptr, BFVM_E_ARG, "null handle");
    Env& e = *h;
    if (e.S > kNoAllocMaxS)
        return rfail(h, BFVM_E_ARG,
                     "bfvm_mass_balance_noalloc: " + std::to_string(e.S) + " substrates exceed its static "
                     "workspace's limit of " + std::to_string(kNoAllocMaxS) + " (use bfvm_mass_balance)");
    void* ws = nullptr;
    hipError_t st = scratch_address(e.device, &ws);
    if (st != hipSuccess) return rhip(h, st, "bfvm_mass_balance_noalloc: workspace address");
    const bfvm::MassLayout m = bfvm::mass_layout(e, ws);
    if (m.bytes > kScratchBytes) return rfail(h, BFVM_E_ARG, "bfvm_mass_balance_noalloc: workspace too small");
    // a member of an ensemble: its store may still be written by the ensemble's stream
    if (e.ensemble_stream && (st = hipStreamSynchronize(e.ensemble_stream)) != hipSuccess)
        return rhip(h, st, "bfvm_mass_balance_noalloc");
    if ((st = bfvm::launch_mass_balance(e, m, sums ? sums : m.sums, nonfinite ? nonfinite : m.count)) !=
        hipSuccess)
        return rhip(h, st, "bfvm_mass_balance_noalloc");
    return BFVM_OK;
}
