// Profiling build of the same kernels: thread 0 of every workgroup adds a timestamp
// (clock64, shader cycles) to one accumulator per phase mark of the fused step.  The
// difference of consecutive accumulators divided by the number of block-iterations is the
// mean cycle count of a phase.  Built as libkynema_amd_phases.so; the library proper
// compiles the marks to nothing.
#include <hip/hip_runtime.h>

__device__ unsigned long long kyn_phase_acc[16];
#define KYN_PHASE_MARK(i)                                                                  \
    do {                                                                                   \
        if (threadIdx.x == 0) atomicAdd(&kyn_phase_acc[i], (unsigned long long)clock64()); \
    } while (0)

__device__ unsigned long long kyn_span_acc[12];
#define KYN_SPAN_BEGIN(v) const unsigned long long v = clock64()
#define KYN_SPAN_END(slot, v, who)                                                              \
    do {                                                                                        \
        if (threadIdx.x == (who)) atomicAdd(&kyn_span_acc[slot], (unsigned long long)(clock64() - (v))); \
    } while (0)

#include "kyn_kernels.hip"

extern "C" int kyn_phase_read(unsigned long long* out) {
    if (hipMemcpyFromSymbol(out, HIP_SYMBOL(kyn_phase_acc), sizeof(unsigned long long) * 16) != hipSuccess) return -1;
    return hipMemcpyFromSymbol(out + 16, HIP_SYMBOL(kyn_span_acc), sizeof(unsigned long long) * 12) == hipSuccess ? 0 : -1;
}
extern "C" int kyn_phase_reset() {
    unsigned long long z[16] = {0};
    if (hipMemcpyToSymbol(HIP_SYMBOL(kyn_phase_acc), z, sizeof(z)) != hipSuccess) return -1;
    return hipMemcpyToSymbol(HIP_SYMBOL(kyn_span_acc), z, sizeof(unsigned long long) * 12) == hipSuccess ? 0 : -1;
}
