// noalloc.hip — gsr_sort_pairs without any device allocation.
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc) get a variant of
// gsr_sort_pairs whose scratch (frame counters, digit histograms and tickets,
// the look-back words of two passes) lives in one fixed device array per GPU,
// noalloc_detail::scratch below, instead of context buffers that grow on
// demand.  The array's name is a host-side stand-in in host code: its device
// address on each GPU comes
// from hipGetSymbolAddress at context creation (noalloc_attach) and only that
// address is handed to kernels and copies.  Every call on one GPU shares the
// array; the kernels clear what must start at zero (counters, histograms,
// tickets and look-back words) on every call, and calls that overlap on one
// GPU race.
#include <mutex>

#include "gsr_internal.h"

// Layout: FrameCounters (256 B) | histograms and tickets | look-back words.
constexpr size_t kNoAllocScratchBytes = (size_t)GSR_SORT_NOALLOC_SCRATCH_BYTES;
constexpr size_t kNoAllocHistOffset = 256;
constexpr size_t kNoAllocStatusOffset = 32 * 1024;
static_assert(sizeof(gsr::FrameCounters) == kNoAllocHistOffset, "counters head the scratch");

namespace gsr {

// The array has external linkage (inside this namespace): hipcc keeps an
// internal (static) device variable as a local symbol of the code object,
// where hipGetSymbolAddress cannot find it.
namespace noalloc_detail {
__device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];
}

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mutex;
uint8_t* g_noalloc_ptr[kMaxDevices] = {};
}  // namespace

void noalloc_attach(Context* c) {
    if (c->device < 0 || c->device >= kMaxDevices) {
        c->noalloc_message = "device ordinal beyond the no-allocation scratch table";
        return;
    }
    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    if (!g_noalloc_ptr[c->device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_detail::scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();  // do not leave the failure for the caller's next HIP call
            c->noalloc_message = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return;
        }
        g_noalloc_ptr[c->device] = (uint8_t*)p;
    }
    c->noalloc_scratch = g_noalloc_ptr[c->device];
}

}  // namespace gsr

using namespace gsr;

extern "C" {

int gsr_sort_pairs_noalloc(gsr_context* ctx, uint32_t key_bytes, void* keys0_dev, uint32_t* values0_dev,
                           void* keys1_dev, uint32_t* values1_dev, uint64_t n, uint32_t begin_bit, uint32_t end_bit,
                           int* result_index) {
    Context* c = reinterpret_cast<Context*>(ctx);
    if (!c) return fail(nullptr, GSR_ERR_ARGUMENT, "null context");
    if (!keys0_dev || !values0_dev || !keys1_dev || !values1_dev || !result_index)
        return fail(c, GSR_ERR_ARGUMENT, "gsr_sort_pairs_noalloc: null argument");
    if (key_bytes != 4 && key_bytes != 8) return fail(c, GSR_ERR_ARGUMENT, "key_bytes must be 4 or 8");
    if (end_bit > key_bytes * 8 || begin_bit > end_bit)
        return fail(c, GSR_ERR_ARGUMENT, "bad bit range [%u, %u) for %u-byte keys", begin_bit, end_bit, key_bytes);
    if (!c->noalloc_scratch)
        return fail(c, GSR_ERR_DEVICE, "no-allocation scratch unavailable: %s", c->noalloc_message.c_str());
    const SortWords need = sort_scratch_words(c, key_bytes, n ? n : 1, begin_bit, end_bit, 1, 8);
    const size_t hist_avail = (kNoAllocStatusOffset - kNoAllocHistOffset) / 4;
    const size_t status_avail = (kNoAllocScratchBytes - kNoAllocStatusOffset) / 4;
    if (n > kMaxSortKeys || need.hist > hist_avail || need.status > status_avail)
        return fail(c, GSR_ERR_ARGUMENT,
                    "%llu %u-byte keys exceed the no-allocation sort's fixed scratch of %zu bytes "
                    "(look-back words for at most %zu keys)",
                    (unsigned long long)n, key_bytes, kNoAllocScratchBytes,
                    (size_t)(status_avail / (2 * 256)) * sort_default_tile_keys(key_bytes));
    hipSetDevice(c->device);
    FrameCounters* counters = (FrameCounters*)c->noalloc_scratch;
    const SortMem mem{(uint32_t*)(c->noalloc_scratch + kNoAllocHistOffset), hist_avail,
                      (uint32_t*)(c->noalloc_scratch + kNoAllocStatusOffset), status_avail};
    int rc = launch_fill_u32(c, (uint32_t*)counters, 0u, sizeof(FrameCounters) / 4);
    if (!rc)
        rc = launch_sort(c, key_bytes, keys0_dev, values0_dev, keys1_dev, values1_dev, nullptr, n, n ? n : 1,
                         begin_bit, end_bit, 1, counters, result_index, 8, false, &mem);
    if (rc) return rc;
    GSR_HIP(c, hipMemcpyAsync(c->host_counters, counters, sizeof(FrameCounters), hipMemcpyDeviceToHost, c->stream));
    GSR_HIP(c, hipStreamSynchronize(c->stream));
    if (c->host_counters->error) return fail(c, GSR_ERR_KERNEL, "sort look-back wait exceeded its bound");
    return GSR_OK;
}

}  // extern "C"
