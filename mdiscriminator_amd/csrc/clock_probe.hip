// clock_probe.hip -- measurement support for bench.py --first-steps: the
// shader clock seen by one lane, sampled between steps of a workload.
//
// clk_probe: lane 0 of one workgroup counts shader-clock ticks (clock64)
// over a fixed span of the constant-rate wall clock (wall_clock64) and stores
// both deltas at out[2*slot], out[2*slot+1]; ticks / (span / wall rate) is the
// shader clock frequency while the probe ran.
#include <hip/hip_runtime.h>

extern "C" __global__ void clk_probe(unsigned long long* out, unsigned slot, unsigned span) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const unsigned long long w0 = wall_clock64();
    const unsigned long long c0 = clock64();
    unsigned long long w1 = w0;
    while (w1 - w0 < span) w1 = wall_clock64();
    const unsigned long long c1 = clock64();
    out[2 * slot] = c1 - c0;
    out[2 * slot + 1] = w1 - w0;
}

// Launches one probe on `stream`; out must hold 2 * (slot + 1) words.
extern "C" int mdiscr_clk_probe(void* out, unsigned slot, unsigned span, void* stream) {
    hipLaunchKernelGGL(clk_probe, dim3(1), dim3(64), 0, (hipStream_t)stream,
                       static_cast<unsigned long long*>(out), slot, span);
    return (int)hipGetLastError();
}

// Ticks per second of wall_clock64 on device 0, in kHz.
extern "C" int mdiscr_wall_clock_khz(int* khz) {
    return (int)hipDeviceGetAttribute(khz, hipDeviceAttributeWallClockRate, 0);
}
