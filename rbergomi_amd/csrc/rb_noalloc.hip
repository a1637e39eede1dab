// rb_noalloc.hip -- rb_price_noalloc: rb_price_v2 on the fused kernel without any device
// allocation, for callers that cannot allocate (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc).
//
// Every device buffer the fused pricing call needs beyond the caller's inputs and outputs -- the
// request's tables (Toeplitz A fragments, P_j, group and option records), the work-item lists,
// the per-item FP64 option sums, the work-queue counter and the device-side results -- is carved
// from one file-scope device array, `scratch`, of NOALLOC_SCRATCH_BYTES.  Each GPU the library
// runs on holds its own copy of the array from module load until the process ends.  Host code
// never uses the array's name as an address: on the first call on a device, with that device
// current, hipGetSymbolAddress gives the device address, which is kept per device and passed to
// the kernels like any other buffer.  The kernels are those of rb_price_v2 (the same tables, the
// same launches), so the prices are bit-identical to rb_price_v2 with the same configuration.
//
// Every call on a device shares its array, which keeps the last call's contents; the work-queue
// counter is cleared by a kernel before each launch.  Calls are serialised by a process-wide
// mutex held until the call's results are on the host, so threads of one process never overlap
// on the array (calls that overlapped on one device would race on it).
#include <hip/hip_runtime.h>

#include <array>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rbergomi_mi355x.h"
#include "rb_engine.h"
#include "rb_fused_host.h"

namespace {

constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;  // 64 MiB per device
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_noalloc_mu;  // serialises calls; guards g_scratch
std::array<void*, NOALLOC_MAX_DEVICES> g_scratch{};

int fail(int rc, const std::string& msg) {
    rb::set_last_error(msg);
    return rc;
}

}  // namespace

extern "C" int rb_price_noalloc(double xi, const double* H, const double* eta, const double* rho, int par_size,
                                const double* T, const double* K, int op_size, int N, long M, const uint64_t* seed,
                                int seed_size, const rb_config* cfg, void* stream, double* price, double* iv,
                                double* stat, int out_size) {
    using namespace rb;
    set_last_error("");
    if (out_size != par_size * op_size) return fail(RB_ERR_SIZE, "out_size != par_size*op_size");
    if (par_size < 1 || op_size < 1) return fail(RB_ERR_ARG, "par_size and op_size must be positive");
    if (!H || !eta || !rho || !T || !K || !price || !stat || !cfg)
        return fail(RB_ERR_ARG, "null input or output array, or no configuration");
    const int mode = fused_mode_of(cfg->precision);
    if (mode < 0 || cfg->scheme != RB_SCHEME_HYBRID)
        return fail(RB_ERR_ARG, "rb_price_noalloc runs the fused kernel: precision RB_PREC_SPLIT_FP16, "
                                "RB_PREC_SPLIT_BF16 or RB_PREC_FP32_GEMM, hybrid scheme");
    Request r;
    r.xi = xi;
    r.H.assign(H, H + par_size);
    r.eta.assign(eta, eta + par_size);
    r.rho.assign(rho, rho + par_size);
    r.T.assign(T, T + op_size);
    r.K.assign(K, K + op_size);
    r.N = N;
    r.M = M;
    if (seed && seed_size > 0) r.seed.assign(seed, seed + seed_size);
    r.precision = cfg->precision;
    r.comp = cfg->compensator;
    r.method = cfg->method;
    std::string err;
    int rc = validate(r, err);
    if (rc) return fail(rc, err);
    if (N > 1024) return fail(RB_ERR_LIMIT, "rb_price_noalloc covers N <= 1024 (the fused Toeplitz-GEMM kernel)");
    const Layout lay = make_layout(r);
    FusedHostReq fh;
    rc = plan_fused_request(r, lay, mode, fh, err);
    if (rc) return fail(rc, err);
    // every carved buffer is rounded up to 256 bytes: at most 19 buffers
    const size_t need = FusedBatch::footprint({&fh}) + 19 * 256;
    if (need > NOALLOC_SCRATCH_BYTES)
        return fail(RB_ERR_LIMIT, "request needs about " + std::to_string(need >> 20) + " MiB of device scratch, above the " +
                                      std::to_string(NOALLOC_SCRATCH_BYTES >> 20) +
                                      " MiB of rb_price_noalloc (fewer paths or options per call)");

    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= NOALLOC_MAX_DEVICES)
        return fail(RB_ERR_DEVICE, "no HIP device");
    if (!g_scratch[device]) {
        // once per device, with that device current: the device address of this GPU's copy
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(RB_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[device] = p;
    }
    DevArena arena;
    arena.base = static_cast<char*>(g_scratch[device]);
    arena.size = NOALLOC_SCRATCH_BYTES;
    hipStream_t st = static_cast<hipStream_t>(stream);
    FusedBatch B;
    B.arena = &arena;
    rc = B.build({&fh}, mode, device, st, err);
    if (!rc) rc = B.launch(st, err);
    if (!rc) rc = B.results(0, price, iv, stat, st, err);
    if (!rc && hipStreamSynchronize(st) != hipSuccess) {
        err = "fused pricing failed on the device";
        rc = RB_ERR_DEVICE;
    }
    if (rc == ST_ERR_NOMEM) {
        // only the arena can run out here
        err = "request does not fit the " + std::to_string(NOALLOC_SCRATCH_BYTES >> 20) + " MiB scratch of rb_price_noalloc";
        rc = RB_ERR_LIMIT;
    }
    if (rc) return fail(rc, err);
    return RB_OK;
}
