// bench_probe.hip -- VALU issue-rate and copy probes used by bench.py --probe
// (bench tooling, not part of the library).
//
// Each kernel runs 8 independent chains of one instruction kind per lane, 16
// repetitions per loop trip, on a full grid (8 workgroups of 4 waves per CU), so
// the measured lane-operation rate is the issue rate of that instruction:
//   kind 0  v_xor_b32        (full-rate reference)
//   kind 1  v_mul_lo_u32     (32-bit multiply, used twice per hash)
//   kind 2  v_lshl_add_u64   (64-bit add of key and seed)
//   kind 3  v_mul_u32_u24    (24-bit multiply)
//   kind 4  v_xor_b32_sdwa   (word-select xor: the 16-bit fold of the hash)
//   kind 5  v_mad_u32_u24    (24-bit multiply-add: remap plus table offset)
//   kind 6  v_lshlrev_b32_sdwa (shift of a selected word: table address)
// rs_probe_valu returns lane-ops per second; bench.py reports the ratio to kind 0
// as the issue cost of each instruction in full-rate slots.
#include <hip/hip_runtime.h>

#include <stdint.h>

namespace {

template <int KIND>
__global__ __launch_bounds__(256) void k_probe(uint32_t iters, uint32_t b, uint32_t *out) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t a[8];
    uint64_t x[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        a[j] = t * 0x9E3779B9u + j;
        x[j] = ((uint64_t)a[j] << 32) | (a[j] ^ b);
    }
    const uint64_t y = ((uint64_t)b << 32) | (b * 3u);
    for (uint32_t it = 0; it < iters; it++) {
#pragma unroll
        for (int r = 0; r < 16; r++) {
#pragma unroll
            for (int j = 0; j < 8; j++) {
                // inline assembly pins the instruction (no algebraic folding of the chains)
                if (KIND == 0) asm volatile("v_xor_b32 %0, %0, %1" : "+v"(a[j]) : "v"(a[(j + 1) & 7]));
                if (KIND == 1) asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(a[j]) : "s"(b));
                if (KIND == 2) asm volatile("v_lshl_add_u64 %0, %0, 0, %1" : "+v"(x[j]) : "v"(y));
                if (KIND == 3) asm volatile("v_mul_u32_u24 %0, %1, %0" : "+v"(a[j]) : "s"(b));
                if (KIND == 4)
                    asm volatile("v_xor_b32_sdwa %0, %0, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:WORD_1 src1_sel:DWORD"
                                 : "+v"(a[j]) : "v"(a[(j + 1) & 7]));
                if (KIND == 5) asm volatile("v_mad_u32_u24 %0, %0, %1, %2" : "+v"(a[j]) : "s"(b), "v"(a[(j + 1) & 7]));
                if (KIND == 6)
                    asm volatile("v_lshlrev_b32_sdwa %0, %1, %0 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:WORD_1"
                                 : "+v"(a[j]) : "v"(a[(j + 1) & 7]));
            }
        }
    }
    uint32_t acc = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) acc ^= a[j] ^ (uint32_t)x[j] ^ (uint32_t)(x[j] >> 32);
    out[t] = acc;
}

// Plain streaming copy (16 B per thread per trip): the calibration kernel for the
// FETCH_SIZE / WRITE_SIZE counters (known bytes read and written).
__global__ __launch_bounds__(256) void k_probe_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// Clock check: wall clock (s_memrealtime) and shader clock (s_memtime) before and
// after a fixed amount of VALU work, from one lane of every workgroup.
__global__ __launch_bounds__(64) void k_probe_clock(uint32_t iters, unsigned long long *out) {
    const unsigned long long w0 = (unsigned long long)wall_clock64(), c0 = (unsigned long long)clock64();
    uint32_t a = threadIdx.x;
    for (uint32_t i = 0; i < iters; i++) asm volatile("v_mul_lo_u32 %0, %0, %0" : "+v"(a));
    const unsigned long long w1 = (unsigned long long)wall_clock64(), c1 = (unsigned long long)clock64();
    if (threadIdx.x == 0) {
        out[4 * blockIdx.x + 0] = w0;
        out[4 * blockIdx.x + 1] = w1;
        out[4 * blockIdx.x + 2] = c0;
        out[4 * blockIdx.x + 3] = c1 + (a & 1);
    }
}

}  // namespace

// Runs k_probe_clock on `blocks` workgroups; out holds 4 values per workgroup
// (wall start, wall end, shader start, shader end); *khz the runtime's wall-clock rate.
extern "C" int rs_probe_clock(int device, uint32_t blocks, uint32_t iters, unsigned long long *out, int *khz) {
    if (hipSetDevice(device) != hipSuccess) return 2;
    if (hipDeviceGetAttribute(khz, hipDeviceAttributeWallClockRate, device) != hipSuccess) return 2;
    unsigned long long *d = nullptr;
    if (hipMalloc(&d, (size_t)blocks * 32) != hipSuccess) return 7;
    k_probe_clock<<<blocks, 64>>>(iters, d);
    const bool ok = hipDeviceSynchronize() == hipSuccess &&
                    hipMemcpy(out, d, (size_t)blocks * 32, hipMemcpyDeviceToHost) == hipSuccess;
    (void)hipFree(d);
    return ok ? 0 : 2;
}

// Copies `bytes` (a multiple of 16) between two device buffers with k_probe_copy.
extern "C" int rs_probe_copy(int device, uint64_t bytes, double *ms_out) {
    if (hipSetDevice(device) != hipSuccess) return 2;
    uint4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess) return 7;
    if (hipMalloc(&b, bytes) != hipSuccess) {
        (void)hipFree(a);
        return 7;
    }
    (void)hipMemset(a, 1, bytes);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, 0);
    k_probe_copy<<<4096, 256>>>(a, b, bytes / 16);
    (void)hipEventRecord(e1, 0);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    *ms_out = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}

extern "C" int rs_probe_valu(int device, int kind, uint32_t iters, double *lane_ops_per_s, double *ms_out) {
    if (hipSetDevice(device) != hipSuccess) return 2;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return 2;
    const uint32_t blocks = (uint32_t)prop.multiProcessorCount * 8, threads = 256;
    uint32_t *out = nullptr;
    if (hipMalloc(&out, (size_t)blocks * threads * 4) != hipSuccess) return 7;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    float best = 1e30f;
    for (int rep = 0; rep < 3; rep++) {
        (void)hipEventRecord(e0, 0);
        switch (kind) {
        case 0: k_probe<0><<<blocks, threads>>>(iters, 0x2545F491u, out); break;
        case 1: k_probe<1><<<blocks, threads>>>(iters, 0x2545F491u, out); break;
        case 2: k_probe<2><<<blocks, threads>>>(iters, 0x2545F491u, out); break;
        case 3: k_probe<3><<<blocks, threads>>>(iters, 0x0045F491u, out); break;
        case 4: k_probe<4><<<blocks, threads>>>(iters, 0x2545F491u, out); break;
        case 5: k_probe<5><<<blocks, threads>>>(iters, 0x0045F491u, out); break;
        default: k_probe<6><<<blocks, threads>>>(iters, 0x2545F491u, out); break;
        }
        (void)hipEventRecord(e1, 0);
        (void)hipEventSynchronize(e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) best = ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(out);
    const double ops = (double)blocks * threads * iters * 16 * 8;
    *lane_ops_per_s = ops / (best / 1e3);
    *ms_out = best;
    return 0;
}
