// Snapshots of the density store in the reference's MAT v4 layout
// (Microenvironment::write_to_matlab, BioFVM_microenvironment.cpp:557-580;
// header write_matlab4_header, BioFVM_matlab.cpp:515-559).
//
// File: a 20-byte header of five 32-bit words (type 0 = little-endian full
// FP64 matrix, rows = 4 + S, cols = voxels, imaginary 0, name length) and the
// name "multiscale_microenvironment" without a terminating NUL (the
// reference writes exactly name.size() bytes), then one column per voxel:
// centre x, y, z, volume, the S densities, all FP64.
//
// A kernel packs the columns on the device (centres x0 + (i + 0.5) dx with a
// separately rounded multiply and add, volume (dx*dy)*dz, as the mesh builds
// them, BioFVM_mesh.cpp:817-860), the packed matrix goes to pinned host
// memory, and the host writes the file.  The asynchronous form snapshots the
// store at that point of the handle's stream, copies on a second stream
// (default priority) and writes from a host thread while later steps run.
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>

#include "bfvm_internal.h"

using bfvm::Env;

namespace bfvm {

struct MatWriter {
    double* dev = nullptr;   // packed matrix on the device
    double* host = nullptr;  // pinned copy
    size_t bytes = 0;
    hipStream_t copy_stream = nullptr;
    hipEvent_t packed = nullptr;
    std::thread worker;
    bool busy = false;
    int result = BFVM_OK;
    std::string err;
};

}  // namespace bfvm

using bfvm::MatWriter;

namespace {

constexpr int kBlock = 256;
const char kMatName[] = "multiscale_microenvironment";

template <typename T>
__global__ void k_pack_matlab(const T* __restrict__ rho, double* __restrict__ out, int nx, int ny, int nz,
                              int S, double x0, double y0, double z0, double dx, double dy, double dz,
                              double dV)
{
#pragma clang fp contract(off)  // the reference's separate roundings, no FMA
    const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nvox = (size_t)nx * ny * nz;
    if (v >= nvox) return;
    const int i = (int)(v % nx);
    const int j = (int)((v / nx) % ny);
    const int k = (int)(v / ((size_t)nx * ny));
    double* col = out + v * (size_t)(4 + S);
    col[0] = x0 + ((double)i + 0.5) * dx;
    col[1] = y0 + ((double)j + 0.5) * dy;
    col[2] = z0 + ((double)k + 0.5) * dz;
    col[3] = dV;
    for (int s = 0; s < S; ++s) col[4 + s] = (double)rho[v * S + s];
}

int iofail(Env& e, int code, const std::string& msg)
{
    e.err = msg;
    return code;
}

#define IOHIP(e, call)                                                                            \
    do {                                                                                          \
        hipError_t st_ = (call);                                                                  \
        if (st_ != hipSuccess) return iofail(e, BFVM_E_DEVICE, std::string(#call) + ": " +        \
                                                                    bfvm::hip_error_text(st_));      \
    } while (0)

// Writes the header and the packed matrix; returns an error message or "".
std::string write_file(const std::string& filename, int rows, int cols, const double* data, size_t bytes)
{
    FILE* fp = std::fopen(filename.c_str(), "wb");
    if (!fp) return "could not open " + filename;
    const unsigned header[5] = {0u, (unsigned)rows, (unsigned)cols, 0u, (unsigned)(sizeof(kMatName) - 1)};
    bool ok = std::fwrite(header, sizeof(unsigned), 5, fp) == 5;
    ok = ok && std::fwrite(kMatName, 1, sizeof(kMatName) - 1, fp) == sizeof(kMatName) - 1;
    ok = ok && std::fwrite(data, 1, bytes, fp) == bytes;
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? std::string() : "short write to " + filename;
}

int join_writer(Env& e)
{
    MatWriter* w = e.mat;
    if (!w || !w->busy) return BFVM_OK;
    if (w->worker.joinable()) w->worker.join();
    w->busy = false;
    if (w->result != BFVM_OK) {
        const int rc = w->result;
        w->result = BFVM_OK;
        return iofail(e, rc, w->err);
    }
    return BFVM_OK;
}

// Packs the store into the writer's device buffer on the handle's stream.
int pack(Env& e, double x0, double y0, double z0)
{
    if (!e.mat) e.mat = new MatWriter;
    MatWriter& w = *e.mat;
    const size_t bytes = e.nvox * (size_t)(4 + e.S) * sizeof(double);
    if (w.bytes != bytes) {
        if (w.dev) (void)hipFree(w.dev);
        if (w.host) (void)hipHostFree(w.host);
        w.dev = nullptr;
        w.host = nullptr;
        w.bytes = 0;
        IOHIP(e, hipMalloc((void**)&w.dev, bytes));
        IOHIP(e, hipHostMalloc((void**)&w.host, bytes, hipHostMallocDefault));
        w.bytes = bytes;
    }
    if (!w.copy_stream) IOHIP(e, hipStreamCreateWithFlags(&w.copy_stream, hipStreamNonBlocking));
    if (!w.packed) IOHIP(e, hipEventCreateWithFlags(&w.packed, hipEventDisableTiming));
    const double dV = (e.dx * e.dy) * e.dz;  // BioFVM_mesh.cpp:831
    const unsigned grid = (unsigned)((e.nvox + kBlock - 1) / kBlock);
    if (e.precision == BFVM_FP64)
        k_pack_matlab<double><<<grid, kBlock, 0, e.stream>>>((const double*)e.rho, w.dev, e.nx, e.ny, e.nz, e.S,
                                                             x0, y0, z0, e.dx, e.dy, e.dz, dV);
    else
        k_pack_matlab<float><<<grid, kBlock, 0, e.stream>>>((const float*)e.rho, w.dev, e.nx, e.ny, e.nz, e.S,
                                                            x0, y0, z0, e.dx, e.dy, e.dz, dV);
    IOHIP(e, hipGetLastError());
    IOHIP(e, hipEventRecord(w.packed, e.stream));
    IOHIP(e, hipStreamWaitEvent(w.copy_stream, w.packed, 0));
    IOHIP(e, hipMemcpyAsync(w.host, w.dev, bytes, hipMemcpyDeviceToHost, w.copy_stream));
    return BFVM_OK;
}

}  // namespace

namespace bfvm {

void free_writer(Env& e)
{
    MatWriter* w = e.mat;
    if (!w) return;
    if (w->worker.joinable()) w->worker.join();
    if (w->copy_stream) (void)hipStreamSynchronize(w->copy_stream);
    if (w->dev) (void)hipFree(w->dev);
    if (w->host) (void)hipHostFree(w->host);
    if (w->packed) (void)hipEventDestroy(w->packed);
    if (w->copy_stream) (void)hipStreamDestroy(w->copy_stream);
    delete w;
    e.mat = nullptr;
}

}  // namespace bfvm

extern "C" {

int bfvm_write_matlab(bfvm_handle h, const char* filename, double x0, double y0, double z0)
{
    bfvm::RandGuard rand_guard;
    if (!h) return BFVM_E_ARG;
    Env& e = *h;
    if (!filename) return iofail(e, BFVM_E_ARG, "filename is NULL");
    int rc;
    if ((rc = join_writer(e))) return rc;
    if (e.ensemble_stream) IOHIP(e, hipStreamSynchronize(e.ensemble_stream));
    if ((rc = pack(e, x0, y0, z0))) return rc;
    IOHIP(e, hipStreamSynchronize(e.mat->copy_stream));
    const std::string msg = write_file(filename, 4 + e.S, (int)e.nvox, e.mat->host, e.mat->bytes);
    if (!msg.empty()) return iofail(e, BFVM_E_ARG, msg);
    return BFVM_OK;
}

int bfvm_write_matlab_async(bfvm_handle h, const char* filename, double x0, double y0, double z0)
{
    bfvm::RandGuard rand_guard;
    if (!h) return BFVM_E_ARG;
    Env& e = *h;
    if (!filename) return iofail(e, BFVM_E_ARG, "filename is NULL");
    int rc;
    if ((rc = join_writer(e))) return rc;  // one snapshot in flight
    if (e.ensemble_stream) IOHIP(e, hipStreamSynchronize(e.ensemble_stream));
    if ((rc = pack(e, x0, y0, z0))) return rc;
    MatWriter& w = *e.mat;
    w.busy = true;
    w.result = BFVM_OK;
    const std::string name(filename);
    const int rows = 4 + e.S, cols = (int)e.nvox;
    w.worker = std::thread([&w, name, rows, cols] {
        const hipError_t st = hipStreamSynchronize(w.copy_stream);
        if (st != hipSuccess) {
            w.result = BFVM_E_DEVICE;
            w.err = std::string("snapshot copy: ") + bfvm::hip_error_text(st);
            return;
        }
        const std::string msg = write_file(name, rows, cols, w.host, w.bytes);
        if (!msg.empty()) {
            w.result = BFVM_E_ARG;
            w.err = msg;
        }
    });
    return BFVM_OK;
}

int bfvm_io_wait(bfvm_handle h)
{
    bfvm::RandGuard rand_guard;
    if (!h) return BFVM_E_ARG;
    return join_writer(*h);
}

}  // extern "C"
