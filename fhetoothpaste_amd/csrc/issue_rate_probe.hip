// Issue-rate probe for the VALU instructions the Goldilocks arithmetic uses on gfx950: each
// kernel runs 8 independent dependency chains of one instruction (or one fixed pair) per
// iteration; lane-ops/clk/CU near 64 means full rate.  Pure ALU work, no memory access
// besides one in-bounds store per thread.  Built by fhetoothpaste_amd/build.py.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdint.h>
// Issue-rate probes: 8 independent dependency chains of one instruction per iteration.
#define CH8(X) X(c0) X(c1) X(c2) X(c3) X(c4) X(c5) X(c6) X(c7)
#define MAD(c) asm volatile("v_mad_u64_u32 %0, vcc, %1, %2, %0" : "+v"(c) : "v"(a), "v"(b) : "vcc");
#define ADDC(c) asm volatile("v_add_co_u32 %0, vcc, %0, %1" : "+v"(c) : "v"(a) : "vcc");
#define MULLO(c) asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(c) : "v"(a));
#define MULHI(c) asm volatile("v_mul_hi_u32 %0, %0, %1" : "+v"(c) : "v"(a));
#define LSHLADD(c) asm volatile("v_lshl_add_u64 %0, %0, 0, %1" : "+v"(c) : "v"(w));
#define CMP64(c) asm volatile("v_cmp_lt_u64 vcc, %0, %1\n v_cndmask_b32 %0, %0, %2, vcc" : "+v"(c32) : "v"(w), "v"(a) : "vcc");
#define CND(c) asm volatile("v_cndmask_b32 %0, %0, %1, vcc" : "+v"(c) : "v"(a) : "vcc");
#define ADD64C(c) asm volatile("v_add_co_u32 %0, vcc, %0, %1\n v_addc_co_u32 %2, vcc, %2, %1, vcc" : "+v"(c), "+v"(a) , "+v"(c32) : : "vcc");
#define SHR64(c) asm volatile("v_lshrrev_b64 %0, 3, %0" : "+v"(c));
#define CMPCND(c) asm volatile("v_cmp_lt_u64 vcc, %0, %1\n v_cndmask_b32 %2, %2, %3, vcc" : "+v"(c), "+v"(w), "+v"(c32) : "v"(a) : "vcc");
#define MUL24(c) asm volatile("v_mul_u32_u24 %0, %0, %1" : "+v"(c) : "v"(a));

#define KERN(NAME, T, BODY) \
__global__ void NAME(uint64_t* out, uint32_t s, int iters) { \
  uint32_t a = threadIdx.x + s, b = blockIdx.x * 7 + 3; uint64_t w = a * 3ull; uint32_t c32 = a; (void)c32; (void)w; (void)b; \
  T c0=a,c1=a+1,c2=a+2,c3=a+3,c4=a+4,c5=a+5,c6=a+6,c7=a+7; \
  for (int i = 0; i < iters; ++i) { CH8(BODY) } \
  out[blockIdx.x * blockDim.x + threadIdx.x] = (uint64_t)(c0^c1^c2^c3^c4^c5^c6^c7); }
KERN(k_mad, uint64_t, MAD)
KERN(k_addc, uint32_t, ADDC)
KERN(k_mullo, uint32_t, MULLO)
KERN(k_mulhi, uint32_t, MULHI)
KERN(k_lshladd, uint64_t, LSHLADD)
KERN(k_cnd, uint32_t, CND)
KERN(k_mul24, uint32_t, MUL24)
KERN(k_add64c, uint32_t, ADD64C)
KERN(k_shr64, uint64_t, SHR64)
KERN(k_cmpcnd, uint64_t, CMPCND)

int main() {
  const int blocks = 256 * 8 * 4, threads = 256, iters = 4000;
  uint64_t* d; (void)hipMalloc(&d, (size_t)blocks * threads * 8);
  hipEvent_t e0, e1; (void)hipEventCreate(&e0); (void)hipEventCreate(&e1);
  struct { const char* name; void (*f)(uint64_t*, uint32_t, int); } ks[] = {
    {"v_mad_u64_u32", k_mad}, {"v_add_co_u32", k_addc}, {"v_mul_lo_u32", k_mullo}, {"v_mul_hi_u32", k_mulhi},
    {"v_lshl_add_u64", k_lshladd}, {"v_cndmask_b32", k_cnd}, {"v_mul_u32_u24", k_mul24}, {"add_co+addc pair", k_add64c}, {"v_lshrrev_b64", k_shr64}, {"cmp_u64+cndmask", k_cmpcnd}};
  for (auto& k : ks) {
    for (int rep = 0; rep < 2; ++rep) {
      (void)hipEventRecord(e0);
      hipLaunchKernelGGL(k.f, dim3(blocks), dim3(threads), 0, 0, d, 1u, iters);
      (void)hipEventRecord(e1); (void)hipEventSynchronize(e1);
      float ms; (void)hipEventElapsedTime(&ms, e0, e1);
      double lane_ops = (double)blocks * threads * iters * 8;
      if (rep) printf("%-16s %8.3f ms  %8.2f T lane-ops/s  (%.2f lane-ops/clk/CU at 2.4 GHz, 256 CUs)\n", k.name, ms,
                      lane_ops / ms / 1e9, lane_ops / (ms * 1e-3) / 2.4e9 / 256);
    }
  }
  (void)hipFree(d);
  return 0;
}
