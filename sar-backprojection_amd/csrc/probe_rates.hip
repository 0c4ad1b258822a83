// Issue-rate probe for gfx950: how many quad-cycles (4 clocks) of a SIMD one wave64 instruction of
// each kind occupies, and the LDS read throughput per compute unit.  A measurement tool (make probe),
// not part of the focusing library: its table and the rocprofv3 counters read on it calibrate the
// VALU-busy figure of the kernels (SQ_ACTIVE_INST_VALU counts quad-cycles of VALU issue per
// instruction, and SQ_ACTIVE_INST_VALU2 the quad-cycles in which two VALU instructions of one wave
// issued together, so busy quad-cycles = ACTIVE_INST_VALU - ACTIVE_INST_VALU2).
//
// Each VALU kernel runs NIT iterations of 8 independent chains of one instruction (inline asm on
// registers only: no memory instruction is written in asm) in every thread; the grid holds 32
// wavefronts per SIMD.  Rate = wave instructions per second / (1024 SIMDs x clock / 4).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace {

constexpr int NIT = 2048;
constexpr double CLOCK_HZ = 2.4e9;  // MI355X peak engine clock
constexpr int SIMDS = 1024;         // 256 CUs x 4

#define CHK(x)                                                                  \
  do {                                                                          \
    hipError_t e_ = (x);                                                        \
    if (e_ != hipSuccess) {                                                     \
      std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      std::exit(1);                                                             \
    }                                                                           \
  } while (0)

#define CHAINS8(T, INSTR)                                                                 \
  __global__ __launch_bounds__(256) void k_##T(double *buf) {                             \
    const int i = blockIdx.x * blockDim.x + threadIdx.x;                                  \
    VT a0 = (VT)buf[i] + 1, a1 = a0 + 1, a2 = a0 + 2, a3 = a0 + 3, a4 = a0 + 4, a5 = a0 + 5, \
       a6 = a0 + 6, a7 = a0 + 7;                                                          \
    for (int k = 0; k < NIT; ++k) {                                                       \
      asm volatile(INSTR : "+v"(a0));                                                     \
      asm volatile(INSTR : "+v"(a1));                                                     \
      asm volatile(INSTR : "+v"(a2));                                                     \
      asm volatile(INSTR : "+v"(a3));                                                     \
      asm volatile(INSTR : "+v"(a4));                                                     \
      asm volatile(INSTR : "+v"(a5));                                                     \
      asm volatile(INSTR : "+v"(a6));                                                     \
      asm volatile(INSTR : "+v"(a7));                                                     \
    }                                                                                     \
    buf[i] = (double)(a0 + a1 + a2 + a3 + a4 + a5 + a6 + a7);                             \
  }

#define VT float
CHAINS8(fma_f32, "v_fma_f32 %0, %0, %0, %0")
CHAINS8(add_f32, "v_add_f32 %0, %0, 1.0")
CHAINS8(mul_f32, "v_mul_f32 %0, %0, 0.5")
CHAINS8(sub_abs_f32, "v_sub_f32_e64 %0, |%0|, 0.5")
CHAINS8(or_b32, "v_or_b32 %0, %0, 7")
CHAINS8(add_u32, "v_add_u32 %0, 7, %0")
CHAINS8(max_f32, "v_max_f32 %0, %0, 1.0")
CHAINS8(cndmask_b32, "v_cndmask_b32 %0, %0, 1.0, vcc")
CHAINS8(rndne_f32, "v_rndne_f32 %0, %0")
CHAINS8(fract_f32, "v_fract_f32 %0, %0")
CHAINS8(cvt_i32_f32, "v_cvt_i32_f32 %0, %0")
CHAINS8(lshl_add_u32, "v_lshl_add_u32 %0, %0, 3, %0")
CHAINS8(sin_f32, "v_sin_f32 %0, %0")
CHAINS8(cos_f32, "v_cos_f32 %0, %0")
#undef VT
#define VT double
CHAINS8(fma_f64, "v_fma_f64 %0, %0, %0, %0")
CHAINS8(add_f64, "v_add_f64 %0, %0, 1.0")
CHAINS8(mul_f64, "v_mul_f64 %0, %0, 0.5")
CHAINS8(rndne_f64, "v_rndne_f64 %0, %0")
CHAINS8(rsq_f64, "v_rsq_f64 %0, %0")
#undef VT

// f32 and f64 multiply-adds alternating in one wave: two f32 instructions issue together only
// when they are adjacent
__global__ __launch_bounds__(256) void k_alt_f32_f64(double *buf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float f0 = (float)buf[i], f1 = f0 + 1, f2 = f0 + 2, f3 = f0 + 3;
  double d0 = buf[i], d1 = d0 + 1, d2 = d0 + 2, d3 = d0 + 3;
  for (int k = 0; k < NIT; ++k) {
    asm volatile("v_fma_f32 %0, %0, %0, %0" : "+v"(f0));
    asm volatile("v_fma_f64 %0, %0, %0, %0" : "+v"(d0));
    asm volatile("v_fma_f32 %0, %0, %0, %0" : "+v"(f1));
    asm volatile("v_fma_f64 %0, %0, %0, %0" : "+v"(d1));
    asm volatile("v_fma_f32 %0, %0, %0, %0" : "+v"(f2));
    asm volatile("v_fma_f64 %0, %0, %0, %0" : "+v"(d2));
    asm volatile("v_fma_f32 %0, %0, %0, %0" : "+v"(f3));
    asm volatile("v_fma_f64 %0, %0, %0, %0" : "+v"(d3));
  }
  buf[i] = f0 + f1 + f2 + f3 + d0 + d1 + d2 + d3;
}

// CH independent chains of v_fma_f32 per thread; the dynamic LDS of the launch limits the
// wavefronts per SIMD
template <int CH>
__global__ __launch_bounds__(256) void k_chains_f32(double *buf) {
  extern __shared__ float lds_pad[];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float a[CH];
#pragma unroll
  for (int c = 0; c < CH; ++c) a[c] = (float)buf[i] + c;
  for (int k = 0; k < NIT; ++k) {
#pragma unroll
    for (int u = 0; u < 8 / CH; ++u)
#pragma unroll
      for (int c = 0; c < CH; ++c) asm volatile("v_fma_f32 %0, %0, %0, %0" : "+v"(a[c]));
  }
  float s = 0;
#pragma unroll
  for (int c = 0; c < CH; ++c) s += a[c];
  if (s == 12345.f) lds_pad[threadIdx.x] = s;  // keeps the LDS allocation; not taken for these inputs
  buf[i] = s;
}

// LDS reads: MODE 0 all lanes one address, 1 four addresses (16 lanes each, adjacent 16-byte
// records), 2 consecutive addresses, 3 pseudo-random addresses; W floats per lane
template <int MODE, int W>
__global__ __launch_bounds__(256) void k_lds(float *out) {
  __shared__ float4 tab[1024];
  for (int i = threadIdx.x; i < 1024; i += 256) tab[i] = make_float4(i, i + 1, i + 2, i + 3);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  float4 acc = make_float4(0, 0, 0, 0);
  unsigned h = lane * 2654435761u;
  for (int k = 0; k < 4096; ++k) {
    int idx;
    if (MODE == 0) {
      idx = k & 1023;
    } else if (MODE == 1) {
      idx = (k * 4 + (lane >> 4)) & 1023;
    } else if (MODE == 2) {
      idx = (k * 64 + lane) & 1023;
    } else {
      h = h * 1664525u + 1013904223u;
      idx = (h >> 8) & 1023;
    }
    if (W == 4) {
      const float4 v = tab[idx];
      acc.x += v.x;
      acc.y += v.y;
      acc.z += v.z;
      acc.w += v.w;
    } else {
      const float2 v = reinterpret_cast<const float2 *>(tab)[idx];
      acc.x += v.x;
      acc.y += v.y;
    }
  }
  out[blockIdx.x * 256 + threadIdx.x] = acc.x + acc.y + acc.z + acc.w;
}

hipEvent_t e0, e1;

template <typename F, typename P>
double time_ms(F f, int blocks, size_t lds, P *buf) {
  for (int r = 0; r < 2; ++r) hipLaunchKernelGGL(f, dim3(blocks), dim3(256), lds, 0, buf);
  CHK(hipGetLastError());
  CHK(hipEventRecord(e0, 0));
  for (int r = 0; r < 5; ++r) hipLaunchKernelGGL(f, dim3(blocks), dim3(256), lds, 0, buf);
  CHK(hipEventRecord(e1, 0));
  CHK(hipEventSynchronize(e1));
  float ms = 0.f;
  CHK(hipEventElapsedTime(&ms, e0, e1));
  return ms / 5;
}

}  // namespace

int main() {
  const int blocks = 8192;  // 32768 wavefronts: 32 per SIMD
  double *buf = nullptr;
  float *fbuf = nullptr;
  CHK(hipMalloc(&buf, (size_t)blocks * 256 * sizeof(double)));
  CHK(hipMemset(buf, 0, (size_t)blocks * 256 * sizeof(double)));
  CHK(hipMalloc(&fbuf, (size_t)blocks * 256 * sizeof(float)));
  CHK(hipEventCreate(&e0));
  CHK(hipEventCreate(&e1));
  const double waves = blocks * 4.0;
  const double quad_per_s = SIMDS * CLOCK_HZ / 4;
  std::printf("| instruction (8 independent chains per wave) | ms | wave instructions per SIMD quad-cycle |\n|---|---|---|\n");
  struct K {
    const char *n;
    void (*f)(double *);
  } ks[] = {{"v_fma_f32", k_fma_f32},       {"v_add_f32", k_add_f32},
            {"v_mul_f32", k_mul_f32},       {"v_sub_f32 (|x| operand)", k_sub_abs_f32},
            {"v_or_b32", k_or_b32},         {"v_add_u32", k_add_u32},
            {"v_max_f32", k_max_f32},       {"v_cndmask_b32", k_cndmask_b32},
            {"v_rndne_f32", k_rndne_f32},   {"v_fract_f32", k_fract_f32},
            {"v_cvt_i32_f32", k_cvt_i32_f32}, {"v_lshl_add_u32", k_lshl_add_u32},
            {"v_sin_f32", k_sin_f32},       {"v_cos_f32", k_cos_f32},
            {"v_fma_f64", k_fma_f64},       {"v_add_f64", k_add_f64},
            {"v_mul_f64", k_mul_f64},       {"v_rndne_f64", k_rndne_f64},
            {"v_rsq_f64", k_rsq_f64},       {"v_fma_f32 / v_fma_f64 alternating", k_alt_f32_f64}};
  for (const K &k : ks) {
    const double ms = time_ms(k.f, blocks, 0, buf);
    std::printf("| %s | %.3f | %.3f |\n", k.n, ms, waves * NIT * 8 / (ms * 1e-3) / quad_per_s);
    std::fflush(stdout);
  }
  // independent chains x wavefronts per SIMD (LDS-limited: 2048 blocks of 4 wavefronts)
  const int W[] = {1, 2, 3, 4, 5, 8};
  const size_t L[] = {96 * 1024, 64 * 1024, 48 * 1024, 40 * 1024, 32 * 1024, 16 * 1024};
  void (*cf[])(double *) = {k_chains_f32<1>, k_chains_f32<2>, k_chains_f32<4>, k_chains_f32<8>};
  const int cn[] = {1, 2, 4, 8};
  for (auto f : cf) CHK(hipFuncSetAttribute((const void *)f, hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024));
  std::printf("\n| v_fma_f32 chains per wave \\ wavefronts per SIMD |");
  for (int w : W) std::printf(" %d |", w);
  std::printf("\n|---|---|---|---|---|---|---|\n");
  for (int c = 0; c < 4; ++c) {
    std::printf("| %d |", cn[c]);
    for (int j = 0; j < 6; ++j) {
      const double ms = time_ms(cf[c], 2048, L[j], buf);
      std::printf(" %.3f |", 2048 * 4.0 * NIT * 8 / (ms * 1e-3) / quad_per_s);
    }
    std::printf("\n");
    std::fflush(stdout);
  }
  std::printf("\n| LDS read | ms | instructions per CU per clock | bytes per CU per clock |\n|---|---|---|---|\n");
  struct L2 {
    const char *n;
    void (*f)(float *);
    int bytes;
  } ls[] = {{"ds_read_b128, one address", k_lds<0, 4>, 16},   {"ds_read_b128, four addresses", k_lds<1, 4>, 16},
            {"ds_read_b128, consecutive", k_lds<2, 4>, 16},   {"ds_read_b128, random", k_lds<3, 4>, 16},
            {"ds_read_b64, consecutive", k_lds<2, 2>, 8},     {"ds_read_b64, random", k_lds<3, 2>, 8}};
  for (const L2 &l : ls) {
    const double ms = time_ms(l.f, blocks, 0, fbuf);
    const double ipc = waves * 4096 / (ms * 1e-3 * CLOCK_HZ * 256);
    std::printf("| %s | %.3f | %.3f | %.0f |\n", l.n, ms, ipc, ipc * 64 * l.bytes);
    std::fflush(stdout);
  }
  CHK(hipFree(buf));
  CHK(hipFree(fbuf));
  return 0;
}
