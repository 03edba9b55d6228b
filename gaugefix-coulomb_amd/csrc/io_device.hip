// io_device.hip -- configuration and gauge-transformation files streamed
// between disk and HBM.
//
// The file formats are those of host_io.c (src/fields_io.c of the reference):
// a configuration is the raw image of 4V matrices of 18 reals of 4 or 8
// bytes, optionally byte-swapped; a gauge transformation is the native-endian
// FP64 image of V matrices.  Here a file moves through two pinned host
// blocks: host threads read (or write) one block with pread (pwrite) while the
// DMA engine copies the other, and the byte swap and the FP32 <-> FP64
// conversions run on the device.  No full-size host image is made.
#include <fcntl.h>
#include <hip/hip_runtime.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

#include "internal.hpp"

using gfc::cplx;
using gfc::kBlock;
using gfcx::fail;
using gfcx::nblocks;

namespace {

constexpr size_t kIoBlock = 16u << 20;  // bytes per pinned block
constexpr int kIoThreads = 4;           // host threads per block

__global__ __launch_bounds__(kBlock) void k_bswap64(uint64_t *__restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = __builtin_bswap64(p[i]);
}

// FP32 reals (optionally byte-swapped) -> FP64.
__global__ __launch_bounds__(kBlock) void k_f32raw_to_f64(const uint32_t *__restrict__ in,
                                                          double *__restrict__ out, int64_t n,
                                                          int swap) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t v = in[i];
    if (swap) v = __builtin_bswap32(v);
    out[i] = (double)__uint_as_float(v);
}

// FP64 reals [0, n) -> out as FP64 or FP32, optionally byte-swapped.
__global__ __launch_bounds__(kBlock) void k_f64_to_raw(const double *__restrict__ in,
                                                       void *__restrict__ out, int64_t n,
                                                       int prec, int swap) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    if (prec == 8) {
        uint64_t v = (uint64_t)__double_as_longlong(in[i]);
        static_cast<uint64_t *>(out)[i] = swap ? __builtin_bswap64(v) : v;
    } else {
        uint32_t v = __float_as_uint((float)in[i]);
        static_cast<uint32_t *>(out)[i] = swap ? __builtin_bswap32(v) : v;
    }
}

// pread / pwrite of n bytes at off, split over kIoThreads threads; false on a
// short transfer or an error.
bool par_io(int fd, char *buf, size_t n, off_t off, bool write) {
    std::vector<std::thread> th;
    std::vector<char> ok(kIoThreads, 1);
    const size_t part = (n + kIoThreads - 1) / kIoThreads;
    for (int t = 0; t < kIoThreads; t++) {
        const size_t a = std::min(n, t * part), b = std::min(n, a + part);
        if (a >= b) break;
        th.emplace_back([=, &ok] {
            size_t done = a;
            while (done < b) {
                const ssize_t r = write ? pwrite(fd, buf + done, b - done, off + done)
                                        : pread(fd, buf + done, b - done, off + done);
                if (r <= 0) {
                    ok[t] = 0;
                    return;
                }
                done += (size_t)r;
            }
        });
    }
    for (auto &x : th) x.join();
    for (char c : ok)
        if (!c) return false;
    return true;
}

// Two pinned blocks of the context, allocated on first use.
int pinned_blocks(gfc_ctx *c) {
    if (c->pin[0]) return GFC_OK;
    for (int k = 0; k < 2; k++) {
        HIP_TRY(hipHostMalloc(&c->pin[k], kIoBlock, hipHostMallocDefault));
        HIP_TRY(hipEventCreateWithFlags(&c->pin_ev[k], hipEventDisableTiming));
    }
    return GFC_OK;
}

// File bytes [0, total) -> device bytes dst[0, total), block by block.
int stream_in(gfc_ctx *c, int fd, char *dst, size_t total) {
    RC_TRY(pinned_blocks(c));
    int k = 0;
    for (size_t off = 0; off < total; off += kIoBlock, k ^= 1) {
        const size_t n = std::min(kIoBlock, total - off);
        HIP_TRY(hipEventSynchronize(c->pin_ev[k]));  // the copy that used this block
        if (!par_io(fd, static_cast<char *>(c->pin[k]), n, (off_t)off, false)) {
            hipStreamSynchronize(c->stream);
            return -2;
        }
        HIP_TRY(hipMemcpyAsync(dst + off, c->pin[k], n, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipEventRecord(c->pin_ev[k], c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return GFC_OK;
}

// Device bytes src[0, total) -> file bytes [0, total), block by block: the
// copy of block b+1 runs while block b is written.
int stream_out(gfc_ctx *c, int fd, const char *src, size_t total) {
    RC_TRY(pinned_blocks(c));
    const size_t nb = (total + kIoBlock - 1) / kIoBlock;
    auto issue = [&](size_t b) -> int {
        const size_t off = b * kIoBlock, n = std::min(kIoBlock, total - off);
        HIP_TRY(hipMemcpyAsync(c->pin[b & 1], src + off, n, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipEventRecord(c->pin_ev[b & 1], c->stream));
        return GFC_OK;
    };
    if (nb) RC_TRY(issue(0));
    for (size_t b = 0; b < nb; b++) {
        HIP_TRY(hipEventSynchronize(c->pin_ev[b & 1]));
        if (b + 1 < nb) RC_TRY(issue(b + 1));
        const size_t off = b * kIoBlock, n = std::min(kIoBlock, total - off);
        if (!par_io(fd, static_cast<char *>(c->pin[b & 1]), n, (off_t)off, true)) {
            hipStreamSynchronize(c->stream);
            return -2;
        }
    }
    return GFC_OK;
}

}  // namespace

extern "C" int gfc_job_load_config(gfc_job *j, const char *path, int in_precision,
                                   int byte_swap) {
    if (in_precision != 4 && in_precision != 8) return -4;
    if (j == nullptr || path == nullptr) return fail(GFC_EINVAL, "gfc_job_load_config: NULL");
    HIP_TRY(hipSetDevice(j->ctx->device));
    const int fd = open(path, O_RDONLY);
    if (fd < 0) {
        fprintf(stderr, "Error: Problem opening config file %s.\n", path);
        return -1;
    }
    struct stat sb;
    const size_t nreal = (size_t)j->V * 4 * 18;
    const size_t total = nreal * (size_t)in_precision;
    if (fstat(fd, &sb) != 0 || (size_t)sb.st_size < total) {
        fprintf(stderr, "Error: Reading from file %s.\n", path);
        close(fd);
        return -2;
    }
    if ((size_t)sb.st_size > total) {
        fprintf(stderr, "Error: File has not been read till the end.\n Check lattice sizes.\n");
        close(fd);
        return -3;
    }
    gfc_ctx *c = j->ctx;
    // FP64 lands in place; FP32 is staged in the u0 fields (288V bytes, the
    // size of an FP32 configuration), which the next fix rebuilds anyway
    char *dst = in_precision == 8 ? reinterpret_cast<char *>(j->cfg)
                                  : reinterpret_cast<char *>(j->u0);
    int rc = stream_in(c, fd, dst, total);
    close(fd);
    if (rc == -2) fprintf(stderr, "Error: Reading from file %s.\n", path);
    if (rc) return rc;
    if (in_precision == 8) {
        if (byte_swap) {
            hipLaunchKernelGGL(k_bswap64, dim3(nblocks(nreal)), dim3(kBlock), 0, c->stream,
                               reinterpret_cast<uint64_t *>(j->cfg), (int64_t)nreal);
            LAUNCH_CHECK();
        }
    } else {
        hipLaunchKernelGGL(k_f32raw_to_f64, dim3(nblocks(nreal)), dim3(kBlock), 0, c->stream,
                           reinterpret_cast<const uint32_t *>(j->u0),
                           reinterpret_cast<double *>(j->cfg), (int64_t)nreal, byte_swap);
        LAUNCH_CHECK();
    }
    RC_TRY(gfcx::set_identity(j));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return GFC_OK;
}

extern "C" int gfc_job_write_gauge_transf(gfc_job *j, const char *path) {
    if (j == nullptr || path == nullptr)
        return fail(GFC_EINVAL, "gfc_job_write_gauge_transf: NULL");
    HIP_TRY(hipSetDevice(j->ctx->device));
    gfc_ctx *c = j->ctx;
    // AoS image of G staged in the u0 fields (144V bytes)
    cplx<double> *stage = j->u0;
    hipLaunchKernelGGL(gfc::k_g_to_aos, dim3(nblocks(j->V)), dim3(kBlock), 0, c->stream, j->g,
                       stage, j->geom.ns, j->Vh);
    LAUNCH_CHECK();
    const int fd = open(path, O_WRONLY | O_CREAT | O_TRUNC, 0644);
    if (fd < 0) {
        hipStreamSynchronize(c->stream);
        return -1;
    }
    int rc = stream_out(c, fd, reinterpret_cast<const char *>(stage),
                        (size_t)j->V * 18 * sizeof(double));
    if (close(fd) != 0 && rc == GFC_OK) rc = -2;
    return rc;
}

extern "C" int gfc_job_write_config(gfc_job *j, const char *path, int out_precision,
                                    int byte_swap) {
    if (out_precision != 4 && out_precision != 8) return -4;
    if (j == nullptr || path == nullptr) return fail(GFC_EINVAL, "gfc_job_write_config: NULL");
    HIP_TRY(hipSetDevice(j->ctx->device));
    gfc_ctx *c = j->ctx;
    const int64_t nreal = j->V * 4 * 18;
    const size_t total = (size_t)nreal * (size_t)out_precision;
    // the output image, staged in the u0 fields (FP32: 288V bytes) or, for
    // FP64, converted in place only when it has to be swapped
    const char *src = reinterpret_cast<const char *>(j->cfg);
    void *tmp = nullptr;
    if (out_precision == 4 || byte_swap) {
        void *stage = j->u0;
        if (out_precision == 8) {
            hipError_t e = hipMalloc(&tmp, total);
            if (e != hipSuccess)
                return fail(GFC_ENOMEM, std::string("gfc_job_write_config: ") +
                                            hipGetErrorString(e));
            stage = tmp;
        }
        hipLaunchKernelGGL(k_f64_to_raw, dim3(nblocks(nreal)), dim3(kBlock), 0, c->stream,
                           reinterpret_cast<const double *>(j->cfg), stage, nreal, out_precision,
                           byte_swap);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            hipFree(tmp);
            return fail(GFC_EDEVICE, hipGetErrorString(e));
        }
        src = static_cast<const char *>(stage);
    }
    const int fd = open(path, O_WRONLY | O_CREAT | O_TRUNC, 0644);
    int rc = -1;
    if (fd >= 0) {
        rc = stream_out(c, fd, src, total);
        if (close(fd) != 0 && rc == GFC_OK) rc = -2;
    } else {
        fprintf(stderr, "Error: Problem creating file %s for config.\n", path);
        hipStreamSynchronize(c->stream);
    }
    hipFree(tmp);
    return rc;
}
