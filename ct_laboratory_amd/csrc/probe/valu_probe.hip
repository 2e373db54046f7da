// valu_probe.hip -- measures the VALU issue rate of one MI355X CU for the
// instruction kinds the traversal is made of: non-packed FP32 fused
// multiply-adds (v_fma_f32), FP32 adds (v_add_f32) and 32-bit integer adds
// (v_add_u32), one instruction per chain step.  Each lane runs eight
// independent dependency chains, so a wave always has an instruction ready;
// launches from one to eight waves per SIMD show how many waves the issue rate
// needs.  The shader clock comes from s_memtime against the 100 MHz
// s_memrealtime counter on the same waves.
//
//   hipcc -O3 --offload-arch=gfx950 -Xclang -target-feature -Xclang -packed-fp32-ops \
//         ct_laboratory_amd/csrc/probe/valu_probe.hip -o /tmp/valu_probe && /tmp/valu_probe
//
// Prints one line per (kind, waves per SIMD): wave instructions per CU per
// shader cycle, lane operations per second, and the clock.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#define CHECK(x)                                                                          \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) {                                                           \
            fprintf(stderr, "%s:%d %s: %s\n", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
            exit(1);                                                                      \
        }                                                                                 \
    } while (0)

enum Kind { FMA = 0, FADD = 1, IADD = 2 };
constexpr int kChains = 8, kUnroll = 8;

// m, c and step come in as arguments (scalar registers), so every chain step is
// one vector instruction with scalar operands (no literal constants).
template <int KIND>
__global__ __launch_bounds__(256) void k_probe(float* out, int iters, unsigned long long* clocks, float m, float c,
                                               unsigned step) {
    const unsigned long long r0 = wall_clock64(), c0 = clock64();
    float f[kChains];
    unsigned u[kChains];
#pragma unroll
    for (int k = 0; k < kChains; ++k) {
        f[k] = 1.0f + 1e-7f * (float)(threadIdx.x + k);
        u[k] = threadIdx.x * 2654435761u + (unsigned)k;
    }
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int j = 0; j < kUnroll; ++j) {
#pragma unroll
            for (int k = 0; k < kChains; ++k) {
                if (KIND == FMA) f[k] = __builtin_fmaf(f[k], m, c);
                if (KIND == FADD) f[k] = f[k] + c;
                if (KIND == IADD) {
                    u[k] = u[k] + step;
                    asm volatile("" : "+v"(u[k]));  // keeps the compiler from folding the adds
                }
            }
        }
    }
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < kChains; ++k) acc += f[k] + (float)u[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = acc;
    const unsigned long long r1 = wall_clock64(), c1 = clock64();
    if (threadIdx.x == 0) {
        clocks[2 * blockIdx.x] = c1 - c0;
        clocks[2 * blockIdx.x + 1] = r1 - r0;
    }
}

template <int KIND>
static void run(const char* name, int cus, int waves_per_simd, float* out, unsigned long long* clocks) {
    const int threads = 256;  // 4 waves per block
    const int blocks = cus * 4 * waves_per_simd / 4;
    const int iters = 4000;
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    const float m = 0.999999f, c = 1e-7f;
    const unsigned step = 0x9e3779b9u;
    hipLaunchKernelGGL(k_probe<KIND>, dim3(blocks), dim3(threads), 0, 0, out, 10, clocks, m, c, step);  // warm
    CHECK(hipEventRecord(e0));
    hipLaunchKernelGGL(k_probe<KIND>, dim3(blocks), dim3(threads), 0, 0, out, iters, clocks, m, c, step);
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    float ms = 0.0f;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    std::vector<unsigned long long> h(2 * blocks);
    CHECK(hipMemcpy(h.data(), clocks, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    double cyc = 0.0, rt = 0.0;
    for (int b = 0; b < blocks; ++b) {
        cyc += (double)h[2 * b];
        rt += (double)h[2 * b + 1];
    }
    const double ghz = (cyc / rt) * 0.1;  // s_memrealtime ticks at 100 MHz
    const double wave_instr = (double)blocks * (threads / 64) * iters * kUnroll * kChains;
    const double per_cu_cycle = wave_instr / ((double)cus * (ms * 1e-3) * ghz * 1e9);
    printf("%-5s waves/SIMD %d: %.3f wave-instr per CU-cycle, %.3e lane-ops/s, clock %.3f GHz, %.3f ms\n", name,
           waves_per_simd, per_cu_cycle, wave_instr * 64.0 / (ms * 1e-3), ghz, ms);
    CHECK(hipEventDestroy(e0));
    CHECK(hipEventDestroy(e1));
}

int main() {
    hipDeviceProp_t p;
    CHECK(hipGetDeviceProperties(&p, 0));
    const int cus = p.multiProcessorCount;
    printf("%s, %d CUs\n", p.gcnArchName, cus);
    float* out;
    unsigned long long* clocks;
    CHECK(hipMalloc(&out, (size_t)cus * 8 * 256 * sizeof(float)));
    CHECK(hipMalloc(&clocks, (size_t)cus * 8 * 2 * sizeof(unsigned long long)));
    for (int w : {1, 2, 4, 8}) {
        run<FMA>("fma", cus, w, out, clocks);
        run<FADD>("fadd", cus, w, out, clocks);
        run<IADD>("iadd", cus, w, out, clocks);
    }
    CHECK(hipFree(out));
    CHECK(hipFree(clocks));
    return 0;
}
