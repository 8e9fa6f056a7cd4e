// noalloc.hip -- smi_sort_exchange_noalloc: the particle exchange and cell
// sort of smi_sort_exchange on a scratch buffer that the library holds in a
// static device array, for callers that cannot allocate device memory.
//
// The array `scratch` (kScratchWords 4-byte words, 64 MiB) exists once per
// GPU, for as long as the process runs.  In host code its name stands for a
// host-side shadow, not a device address: the device address of each GPU's
// copy comes from hipGetSymbolAddress, called once per device with that
// device current, and only that address is handed to kernels and copies.
// Every call on one GPU uses the same array, so calls that overlap on one
// GPU (two host threads, or two streams in flight) race; the sort clears
// what must start at zero (cell counts, placed-stayer flags, error flag)
// itself, so the array may hold anything from the previous call.
#include <mutex>
#include <string>

#include "../../include/smilei_amd.h"
#include "kernels.h"
#include "seams_internal.h"

namespace {
constexpr int64_t kScratchWords = int64_t(16) << 20;  // 64 MiB
}

static __device__ __align__(256) int32_t scratch[kScratchWords];

namespace smi {
namespace {

constexpr int kMaxDevices = 64;

std::mutex g_mtx;
int32_t *g_ptr[kMaxDevices] = {};

// Device address of the current device's copy of `scratch` (looked up once
// per device); nullptr after recording the error.
int32_t *scratch_ptr()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        seams::op_fail(SMI_ERR_HIP, "no usable current device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_mtx);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            seams::op_fail(SMI_ERR_HIP, std::string("hipGetSymbolAddress of the static scratch failed: ") +
                                            hipGetErrorString(e));
            return nullptr;
        }
        g_ptr[dev] = (int32_t *)p;
    }
    return g_ptr[dev];
}

}  // namespace
}  // namespace smi

using namespace smi;

extern "C" {

int64_t smi_sort_exchange_noalloc_max_words(void) { return kScratchWords; }

int smi_sort_exchange_noalloc(int ndim, const int n_cells[3], const int patch_size[3], const double cell_length[3],
                              int64_t n, double *x, double *y, double *z, double *px, double *py, double *pz,
                              double *w, int16_t *charge, int32_t *key, int32_t *patch_count, int32_t *bin_index,
                              void *stream)
{
    const seams::SortArrays A{x, y, z, px, py, pz, w, charge, key, patch_count, bin_index};
    Geom g;
    int rc = seams::sort_exchange_args(ndim, n_cells, patch_size, cell_length, n, A, g);
    if (rc) return rc;
    const int64_t words = seams::sort_scratch_words(g, n);
    if (words > kScratchWords)
        return seams::op_fail(SMI_ERR_CAPACITY, "sort/exchange needs " + std::to_string(words) +
                                                    " words of scratch; the static buffer holds " +
                                                    std::to_string(kScratchWords));
    int32_t *base = scratch_ptr();
    if (!base) return SMI_ERR_HIP;
    seams::SortScratch s = seams::carve_sort_scratch(g, n, base);
    try {
        rc = seams::sort_exchange_impl(g, n, A, s, (hipStream_t)stream);
    } catch (const std::exception &e) {
        rc = seams::op_fail(SMI_ERR_HIP, e.what());
    }
    (void)hipStreamSynchronize((hipStream_t)stream);
    return rc;
}

}  // extern "C"
