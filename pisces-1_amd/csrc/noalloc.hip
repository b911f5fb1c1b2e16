// noalloc.hip -- the kinetic preconditioner without device allocation.
//
// pisces_kinetic_precond allocates its device buffers per call: the residual
// and result copies, the task descriptors and the FP32 intermediate between the
// three plane passes of the split-BF16 form.  Callers that cannot allocate
// (code under stream capture, tight serving loops) pass device residuals and
// results instead and the descriptors and intermediate live in `scratch`, a
// fixed-size array in the library's code object.  Every GPU holds its own copy
// of the array until the process ends; its address on each device is looked up
// once, when the first context is created on that device.
#include <mutex>

#include "context.hpp"
#include "fdm.hpp"

using namespace pisces;

namespace {

constexpr size_t kScratchBytes = size_t(32) << 20;  // 32 MiB
constexpr int kMaxDevices = 64;

// Device-side storage of the allocation-free entry point.  In host code this
// name is a host-side handle: only the per-device address that
// hipGetSymbolAddress returns is passed to kernels or copies.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices];

// One descriptor per vector: the same fields pisces_kinetic_precond fills on
// the host, written on the device so that no host-to-device copy is needed.
__global__ void fdm_tasks_kernel(FdmTask *tasks, int nvec, AxisMats ax0, AxisMats ax1,
                                 AxisMats ax2, const double *r, double *t, uint16_t *work,
                                 long long ngp, double shift, double scale, double floor)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvec) return;
    FdmTask f;
    f.ax[0] = ax0;
    f.ax[1] = ax1;
    f.ax[2] = ax2;
    f.src = r + ngp * v;
    f.dst = t + ngp * v;
    f.work = work ? work + ngp * v : nullptr;
    f.shift = shift;
    f.scale = scale;
    f.floor = floor;
    f.src32 = 0;
    f.dst32 = 0;
    f.absden = 0;
    tasks[v] = f;
}

inline size_t round256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

namespace pisces {

// Address of `scratch` on `device`, which must be the current device.  Looked
// up once per device and cached (the address of a code-object variable does not
// change while the library is loaded).
void *noalloc_scratch(int device)
{
    PC_REQUIRE(device >= 0 && device < kMaxDevices, PISCES_ERR_ARG, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        PC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        PC_REQUIRE(p != nullptr, PISCES_ERR_HIP, "no device address for the scratch array");
        g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}

}  // namespace pisces

extern "C" {

size_t pisces_noalloc_scratch_bytes(void) { return kScratchBytes; }

int pisces_kinetic_precond_noalloc(pisces_ctx *c, int nvec, const double *d_r, double shift,
                                   double scale, double floor, int precision, double *d_t)
{
    if (!c || nvec <= 0 || !d_r || !d_t || !(floor > 0)) return PISCES_ERR_ARG;
    if (precision != PISCES_FP64 && precision != PISCES_BF16X2) return PISCES_ERR_ARG;
    if (d_r == d_t) {
        set_last_error("pisces_kinetic_precond_noalloc: the residual and result buffers must differ");
        return PISCES_ERR_ARG;
    }
    if (!c->scratch) {
        set_last_error("pisces_kinetic_precond_noalloc: no scratch address for this context");
        return PISCES_ERR_STATE;
    }
    const size_t task_bytes = round256(sizeof(FdmTask) * (size_t)nvec);
    const size_t work_bytes =
        precision == PISCES_BF16X2 ? sizeof(uint16_t) * (size_t)c->ngp * (size_t)nvec : 0;
    if (task_bytes + work_bytes > kScratchBytes) {
        set_last_error("pisces_kinetic_precond_noalloc: " + std::to_string(nvec) +
                       " vectors of " + std::to_string(c->ngp) + " points need " +
                       std::to_string(task_bytes + work_bytes) + " scratch bytes, above the " +
                       std::to_string(kScratchBytes) + " the library holds per device");
        return PISCES_ERR_ARG;
    }
    return guarded_call([&] {
        PC_HIP(hipSetDevice(c->device));
        unsigned char *base = static_cast<unsigned char *>(c->scratch);
        FdmTask *tasks = reinterpret_cast<FdmTask *>(base);
        uint16_t *work = work_bytes ? reinterpret_cast<uint16_t *>(base + task_bytes) : nullptr;
        hipLaunchKernelGGL(fdm_tasks_kernel, dim3((nvec + 63) / 64), dim3(64), 0, c->stream,
                           tasks, nvec, c->ax[0].m, c->ax[1].m, c->ax[2].m, d_r, d_t, work,
                           c->ngp, shift, scale, floor);
        PC_HIP(hipGetLastError());
        fdm_launch(tasks, nvec, c->maxn, precision, c->stream);
        c->sync();
        return 0;
    });
}

}  // extern "C"
