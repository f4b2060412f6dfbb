// linac_rref_u32_noalloc: the plain GF(p) method on a device matrix with no
// allocation at all, for callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops, harnesses that forbid it).
//
// Every buffer the method needs beyond the caller's matrix (elimination
// state, pivot lists, row flags, the gather buffer of the final permutation)
// is carved from `scratch`, a static device array of NOALLOC_BYTES that each
// GPU the library runs on holds until the process ends.  Host code never
// passes the array's name to a kernel or a copy: that name is a host-side
// stand-in.  linac_init (or the entry point's first call on a device) asks
// the runtime for the array's address on that device with
// hipGetSymbolAddress and keeps it per device; kernels receive that address.
//
// Every call on a GPU shares that GPU's array, and the array keeps the last
// call's contents: the method clears what must start at zero (its state and
// row flags) on every call, and calls that overlap on one GPU (from two
// threads, or on two streams) race -- the entry point serialises calls from
// one process through the device context's lock, but another stream of the
// caller's that touches the same matrix is not ordered against it.
#include <mutex>

#include "gf_common.hpp"
#include "runtime.hpp"

namespace linac {

constexpr size_t NOALLOC_BYTES = (size_t)32 << 20;

static __device__ __align__(256) uint8_t scratch[NOALLOC_BYTES];

static std::mutex g_noalloc_mu;
static void* g_scratch_dev[64] = {};  // the array's address on each device

int noalloc_setup(int device) {
    if (device < 0 || device >= 64) return fail(LINAC_ERR_DEVICE, "device %d out of range", device);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_scratch_dev[device]) return LINAC_OK;
    int cur = -1;
    LINAC_HIP(hipGetDevice(&cur));
    if (cur != device) return fail(LINAC_ERR_DEVICE, "noalloc set-up for device %d while device %d is current", device, cur);
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        (void)hipGetLastError();
        return fail(LINAC_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    }
    g_scratch_dev[device] = ptr;
    return LINAC_OK;
}

}  // namespace linac

using namespace linac;

extern "C" {

int linac_rref_u32_noalloc(uint32_t* a, int64_t rows, int64_t cols, int64_t ld, uint32_t p, uint32_t flags,
                           void* stream, int64_t* pivots_out, int64_t* rank_out) {
    if (rows < 0 || cols < 0) return fail(LINAC_ERR_INVALID, "negative shape (%lld, %lld)", (long long)rows, (long long)cols);
    if (ld < cols) return fail(LINAC_ERR_INVALID, "ld (%lld) < cols (%lld)", (long long)ld, (long long)cols);
    if (!(flags & LINAC_DEVICE_PTR)) return fail(LINAC_ERR_INVALID, "linac_rref_u32_noalloc needs a device matrix (LINAC_DEVICE_PTR)");
    const uint32_t method = flags & LINAC_METHOD_MASK;
    if (method != LINAC_METHOD_PLAIN && method != LINAC_METHOD_AUTO)
        return fail(LINAC_ERR_INVALID, "linac_rref_u32_noalloc runs the plain method only");
    if (flags & LINAC_PROFILE) return fail(LINAC_ERR_INVALID, "LINAC_PROFILE creates events: not available without allocation");
    if (!is_prime_u64(p)) return fail(LINAC_ERR_INVALID, "field characteristic %u is not prime", p);
    if (rank_out) *rank_out = 0;
    if (rows == 0 || cols == 0) return LINAC_OK;
    if (!a) return fail(LINAC_ERR_INVALID, "matrix pointer is NULL");
    if ((ld * 4) % 16 != 0 || ((uintptr_t)a & 15) != 0)
        return fail(LINAC_ERR_INVALID, "device matrix rows must be 16-byte aligned (ld*4 %% 16 == 0)");
    const size_t need = noalloc_scratch_bytes(rows, cols, ld, sizeof(uint32_t));
    if (need > NOALLOC_BYTES)
        return fail(LINAC_ERR_INVALID,
                    "a %lld x %lld matrix (ld %lld) needs %zu bytes of scratch, above the no-allocation limit of %zu "
                    "(about rows * ld * 4 + 17 * min(rows, cols) bytes)",
                    (long long)rows, (long long)cols, (long long)ld, need, NOALLOC_BYTES);
    DeviceContext* ctx = nullptr;
    LINAC_TRY(existing_context(&ctx));
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    LINAC_TRY(noalloc_setup(ctx->device));
    uint8_t* sp = (uint8_t*)g_scratch_dev[ctx->device];
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    return rref_u32_noalloc(a, rows, cols, ld, p, flags, s, ctx, sp, pivots_out, rank_out);
}

}  // extern "C"
