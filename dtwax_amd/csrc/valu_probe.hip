// valu_probe.hip -- measured VALU issue limits of the GPU the sDTW kernels run on.
//
// The sDTW kernels are bound by VALU issue, not by memory: the exact binary16
// form issues 4 packed ops per (fwd, rev) cell pair, the binary32 form 4 ops
// per cell.  Their roofline prices those ops at the nominal packed-VALU peak
// (N_CU x 64 lanes x 2 x f_clk).  This probe measures what the hardware
// actually sustains for the same instructions: each kernel issues long runs
// of independent VALU ops (sources v40..v55 never written, destinations
// v56..v63 never read), so only the issue rate limits it, with 1..4 waves per
// SIMD.  Rates are wave64 instructions per CU-cycle and element ops per
// CU-cycle at the device's nominal clock, like the roofline.
//
// The timed loops touch no memory; each lane stores one float at the end.
// Usage: valu_probe [iterations]   (default 20000; prints a table and one
// JSON line)
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>

#include <vector>

#define R8(x) x x x x x x x x
#define PROBE_CLOBBER                                                                                       \
  "v40", "v41", "v42", "v43", "v44", "v45", "v46", "v47", "v48", "v49", "v50", "v51", "v52", "v53", "v54", \
      "v55", "v56", "v57", "v58", "v59", "v60", "v61", "v62", "v63"
// 8 independent instructions per body, 64 per loop iteration
#define PROBE_KERNEL(NAME, BODY)                                            \
  __global__ __launch_bounds__(64) void NAME(int iters, float* out) {       \
    for (int i = 0; i < iters; i++) asm volatile(R8(BODY)::: PROBE_CLOBBER); \
    out[blockIdx.x * 64 + threadIdx.x] = 0.0f;                              \
  }
// two- and three-source forms over registers in distinct banks (mod 4)
#define OPS2(op)                                                                                  \
  op " v56, v41, v42\n" op " v57, v45, v46\n" op " v58, v49, v50\n" op " v59, v53, v54\n" op \
     " v60, v43, v44\n" op " v61, v47, v48\n" op " v62, v51, v52\n" op " v63, v55, v40\n"
#define OPS3(op)                                                                                          \
  op " v56, v41, v42, v43\n" op " v57, v45, v46, v47\n" op " v58, v49, v50, v51\n" op             \
     " v59, v53, v54, v55\n" op " v60, v41, v42, v43\n" op " v61, v45, v46, v47\n" op " v62, v49, v50, v51\n" op \
     " v63, v53, v54, v55\n"

PROBE_KERNEL(k_pk_add_f16, OPS2("v_pk_add_f16"))
PROBE_KERNEL(k_pk_min_f16, OPS2("v_pk_min_f16"))
PROBE_KERNEL(k_pk_mul_f16, OPS2("v_pk_mul_f16"))
PROBE_KERNEL(k_pk_fma_f16, OPS3("v_pk_fma_f16"))
PROBE_KERNEL(k_add_f32, OPS2("v_add_f32"))
PROBE_KERNEL(k_min_f32, OPS2("v_min_f32"))
PROBE_KERNEL(k_mul_f32, OPS2("v_mul_f32"))
PROBE_KERNEL(k_fma_f32, OPS3("v_fma_f32"))
PROBE_KERNEL(k_add_f16, OPS2("v_add_f16"))
// the binary32 cell's mix: fma, mul, min, add in turn
PROBE_KERNEL(k_cell_f32, "v_fma_f32 v56, v41, v42, v43\n v_mul_f32 v57, v45, v45\n v_min_f32 v58, v49, v50\n"
                         " v_add_f32 v59, v53, v54\n v_fma_f32 v60, v45, v46, v47\n v_mul_f32 v61, v49, v49\n"
                         " v_min_f32 v62, v51, v52\n v_add_f32 v63, v55, v40\n")
PROBE_KERNEL(k_min_i32, OPS2("v_min_i32"))
// the binary32 cell's mix with the min on the bit patterns (values >= +0)
PROBE_KERNEL(k_cell_f32_imin, "v_fma_f32 v56, v41, v42, v43\n v_mul_f32 v57, v45, v45\n v_min_i32 v58, v49, v50\n"
                              " v_add_f32 v59, v53, v54\n v_fma_f32 v60, v45, v46, v47\n v_mul_f32 v61, v49, v49\n"
                              " v_min_i32 v62, v51, v52\n v_add_f32 v63, v55, v40\n")
// other forms of the binary32 min: do any of them pair like add/mul?
PROBE_KERNEL(k_max_f32, OPS2("v_max_f32"))
PROBE_KERNEL(k_min3_f32, OPS3("v_min3_f32"))
PROBE_KERNEL(k_med3_f32, OPS3("v_med3_f32"))
PROBE_KERNEL(k_minimum3_f32, OPS3("v_minimum3_f32"))
// the 64-bit (VOP3) encoding of an op that pairs in its 32-bit form
PROBE_KERNEL(k_add_f32_e64, OPS2("v_add_f32_e64"))
// a min beside an fma: does the min pair with another kind of op?
PROBE_KERNEL(k_fma_min_alt, "v_fma_f32 v56, v41, v42, v43\n v_min_f32 v57, v45, v46\n v_fma_f32 v58, v49, v50, v51\n"
                            " v_min_f32 v59, v53, v54\n v_fma_f32 v60, v45, v46, v47\n v_min_f32 v61, v49, v50\n"
                            " v_fma_f32 v62, v51, v52, v53\n v_min_f32 v63, v55, v40\n")
// the binary32 cell with every op in the 32-bit encoding (sub for the row cost)
PROBE_KERNEL(k_cell_f32_vop2, "v_sub_f32 v56, v41, v42\n v_mul_f32 v57, v45, v45\n v_min_f32 v58, v49, v50\n"
                              " v_add_f32 v59, v53, v54\n v_sub_f32 v60, v45, v46\n v_mul_f32 v61, v49, v49\n"
                              " v_min_f32 v62, v51, v52\n v_add_f32 v63, v55, v40\n")

// packed binary16 and binary32 ops alternating (do they issue side by side?)
PROBE_KERNEL(k_mix_pk_f32, "v_pk_add_f16 v56, v41, v42\n v_add_f32 v57, v45, v46\n v_pk_add_f16 v58, v49, v50\n"
                           " v_add_f32 v59, v53, v54\n v_pk_add_f16 v60, v43, v44\n v_add_f32 v61, v47, v48\n"
                           " v_pk_add_f16 v62, v51, v52\n v_add_f32 v63, v55, v40\n")

// the lane moves of the sweeps: a DPP wave shift alone, and one beside seven
// packed binary16 ops (what a lane move costs next to the cells)
PROBE_KERNEL(k_dpp, "v_mov_b32_dpp v56, v41 wave_shr:1 row_mask:0xf bank_mask:0xf\n"
                    " v_mov_b32_dpp v57, v45 wave_shl:1 row_mask:0xf bank_mask:0xf\n"
                    " v_mov_b32_dpp v58, v49 wave_shr:1 row_mask:0xf bank_mask:0xf\n"
                    " v_mov_b32_dpp v59, v53 wave_shl:1 row_mask:0xf bank_mask:0xf\n"
                    " v_mov_b32_dpp v60, v43 wave_shr:1 row_mask:0xf bank_mask:0xf\n"
                    " v_mov_b32_dpp v61, v47 wave_shl:1 row_mask:0xf bank_mask:0xf\n"
                    " v_mov_b32_dpp v62, v51 wave_shr:1 row_mask:0xf bank_mask:0xf\n"
                    " v_mov_b32_dpp v63, v55 wave_shl:1 row_mask:0xf bank_mask:0xf\n")
PROBE_KERNEL(k_pk_dpp, "v_pk_add_f16 v56, v41, v42\n v_pk_add_f16 v57, v45, v46\n v_pk_add_f16 v58, v49, v50\n"
                       " v_pk_add_f16 v59, v53, v54\n v_pk_add_f16 v60, v43, v44\n v_pk_add_f16 v61, v47, v48\n"
                       " v_pk_add_f16 v62, v51, v52\n v_mov_b32_dpp v63, v55 wave_shr:1 row_mask:0xf bank_mask:0xf\n")

typedef void (*probe_fn)(int, float*);

#define CHECK(x)                                                              \
  do {                                                                        \
    hipError_t e_ = (x);                                                      \
    if (e_ != hipSuccess) {                                                   \
      fprintf(stderr, "valu_probe: %s: %s\n", #x, hipGetErrorString(e_));     \
      exit(1);                                                                \
    }                                                                         \
  } while (0)

int main(int argc, char** argv) {
  const int iters = argc > 1 ? atoi(argv[1]) : 20000;
  if (iters < 1 || iters > 1000000) {
    fprintf(stderr, "valu_probe: iterations must be 1..1000000\n");
    return 2;
  }
  int ncu = 0, clk_khz = 0;
  CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, 0));
  CHECK(hipDeviceGetAttribute(&clk_khz, hipDeviceAttributeClockRate, 0));
  hipDeviceProp_t prop;
  CHECK(hipGetDeviceProperties(&prop, 0));
  struct Probe {
    const char* name;
    probe_fn f;
    double elems;  // element ops per wave64 instruction lane (packed = 2)
  } probes[] = {
      {"v_pk_add_f16", k_pk_add_f16, 2.0}, {"v_pk_min_f16", k_pk_min_f16, 2.0}, {"v_pk_mul_f16", k_pk_mul_f16, 2.0},
      {"v_pk_fma_f16", k_pk_fma_f16, 2.0}, {"v_add_f32", k_add_f32, 1.0},       {"v_min_f32", k_min_f32, 1.0},
      {"v_mul_f32", k_mul_f32, 1.0},       {"v_fma_f32", k_fma_f32, 1.0},       {"v_add_f16", k_add_f16, 1.0},
      {"v_min_i32", k_min_i32, 1.0},       {"f32_cell_mix_imin", k_cell_f32_imin, 1.0},
      {"f32_cell_mix", k_cell_f32, 1.0},   {"pk_f16+f32_alt", k_mix_pk_f32, 1.5}, {"v_mov_b32_dpp", k_dpp, 1.0},
      {"7_pk_f16+1_dpp", k_pk_dpp, 14.0 / 8.0}, {"v_max_f32", k_max_f32, 1.0}, {"v_min3_f32", k_min3_f32, 1.0},
      {"v_med3_f32", k_med3_f32, 1.0},     {"v_minimum3_f32", k_minimum3_f32, 1.0},
      {"v_add_f32_e64", k_add_f32_e64, 1.0}, {"fma+min_alt", k_fma_min_alt, 1.0},
      {"f32_cell_mix_vop2", k_cell_f32_vop2, 1.0}};
  const int nprobe = (int)(sizeof(probes) / sizeof(probes[0]));
  const int max_waves = 4;
  std::vector<double> rate_v((size_t)max_waves * nprobe);  // rate(w - 1, k): w waves per SIMD, probe k
  auto rate = [&](int w1, int k) -> double& { return rate_v[(size_t)w1 * nprobe + k]; };
  float* out = nullptr;
  CHECK(hipMalloc(&out, (size_t)ncu * 4 * max_waves * 64 * sizeof(float)));
  hipEvent_t a, b;
  CHECK(hipEventCreate(&a));
  CHECK(hipEventCreate(&b));
  printf("# %s, %d CUs, nominal clock %d MHz, %d iterations x 64 instructions per wave\n", prop.gcnArchName, ncu,
         clk_khz / 1000, iters);
  printf("# wave64 instructions per CU-cycle (element ops per CU-cycle = x 64 lanes x elements)\n");
  printf("%-16s", "instruction");
  for (int w = 1; w <= max_waves; w++) printf("  %d wave%s/SIMD", w, w > 1 ? "s" : " ");
  printf("   elements\n");
  for (int k = 0; k < nprobe; k++) {
    printf("%-16s", probes[k].name);
    for (int w = 1; w <= max_waves; w++) {
      const int blocks = ncu * 4 * w;  // w waves on each of a CU's 4 SIMDs
      hipLaunchKernelGGL(probes[k].f, dim3(blocks), dim3(64), 0, 0, 16, out);  // warm-up
      CHECK(hipGetLastError());
      CHECK(hipEventRecord(a, 0));
      hipLaunchKernelGGL(probes[k].f, dim3(blocks), dim3(64), 0, 0, iters, out);
      CHECK(hipGetLastError());
      CHECK(hipEventRecord(b, 0));
      CHECK(hipEventSynchronize(b));
      float ms = 0.0f;
      CHECK(hipEventElapsedTime(&ms, a, b));
      const double instr = (double)blocks * iters * 64.0;
      rate(w - 1, k) = instr / ncu / ((double)ms * 1e-3 * clk_khz * 1e3);
      printf("  %13.3f", rate(w - 1, k));
    }
    printf("   %.1f\n", probes[k].elems);
  }
  printf("{\"device\": \"%s\", \"n_cu\": %d, \"clock_mhz\": %d, \"unit\": \"wave64 instructions per CU-cycle\", "
         "\"rates\": {",
         prop.gcnArchName, ncu, clk_khz / 1000);
  for (int k = 0; k < nprobe; k++) {
    printf("%s\"%s\": [", k ? ", " : "", probes[k].name);
    for (int w = 1; w <= max_waves; w++) printf("%s%.3f", w > 1 ? ", " : "", rate(w - 1, k));
    printf("]");
  }
  printf("}}\n");
  CHECK(hipEventDestroy(a));
  CHECK(hipEventDestroy(b));
  CHECK(hipFree(out));
  return 0;
}
