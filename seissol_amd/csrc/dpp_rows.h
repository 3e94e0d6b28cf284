// Row broadcasts for VALU products whose one factor is uniform over a group
// of 16 lanes: a row register read once (lane c of each 16-lane row holds
// element c of a 16-element run) serves as sixteen uniform operands through
// DPP row_newbcast, where a broadcast read from LDS would return every value
// to all 64 lanes.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace sdg {

// orders the wave's LDS accesses before and after it (the lanes of one wave
// exchange data through LDS)
__device__ inline void waveSync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// compile-time loop: f(std::integral_constant<int, I>) for I = 0..N-1
template <int N, int I = 0, typename F>
__device__ inline void staticFor(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    staticFor<N, I + 1>(f);
  }
}

// acc[i] += (lane J of row[i] within its row of 16 lanes) * x, i = 0..8:
// nine FMAs whose first factor is broadcast across each 16-lane row (DPP
// row_newbcast).  A row register read from LDS once (lane c of each row holds
// element c of a 16-element run) thereby serves as sixteen wave-uniform
// operands, where a broadcast read from LDS would return every value to all
// 64 lanes.  The leading s_nop 1 gives the two wait states a DPP read needs
// after a VALU write of its source register.
template <int J, typename real>
__device__ inline void fmacRowBcast9(real (&acc)[9], const real (&row)[9], real x) {
  static_assert(J >= 0 && J < 16, "row_newbcast selects a lane of a 16-lane row");
  if constexpr (sizeof(real) == 8) {
    asm("s_nop 1\n\t"
        "v_fmac_f64_dpp %0, %9, %18 row_newbcast:%19 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f64_dpp %1, %10, %18 row_newbcast:%19 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f64_dpp %2, %11, %18 row_newbcast:%19 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f64_dpp %3, %12, %18 row_newbcast:%19 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f64_dpp %4, %13, %18 row_newbcast:%19 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f64_dpp %5, %14, %18 row_newbcast:%19 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f64_dpp %6, %15, %18 row_newbcast:%19 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f64_dpp %7, %16, %18 row_newbcast:%19 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f64_dpp %8, %17, %18 row_newbcast:%19 row_mask:0xf bank_mask:0xf"
        : "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]), "+v"(acc[3]), "+v"(acc[4]), "+v"(acc[5]), "+v"(acc[6]),
          "+v"(acc[7]), "+v"(acc[8])
        : "v"(row[0]), "v"(row[1]), "v"(row[2]), "v"(row[3]), "v"(row[4]), "v"(row[5]), "v"(row[6]),
          "v"(row[7]), "v"(row[8]), "v"(x), "i"(J));
  } else {
    asm("s_nop 1\n\t"
        "v_fmac_f32_dpp %0, %9, %18 row_newbcast:%19 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %1, %10, %18 row_newbcast:%19 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %2, %11, %18 row_newbcast:%19 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %3, %12, %18 row_newbcast:%19 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %4, %13, %18 row_newbcast:%19 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %5, %14, %18 row_newbcast:%19 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %6, %15, %18 row_newbcast:%19 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %7, %16, %18 row_newbcast:%19 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %8, %17, %18 row_newbcast:%19 row_mask:0xf bank_mask:0xf"
        : "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]), "+v"(acc[3]), "+v"(acc[4]), "+v"(acc[5]), "+v"(acc[6]),
          "+v"(acc[7]), "+v"(acc[8])
        : "v"(row[0]), "v"(row[1]), "v"(row[2]), "v"(row[3]), "v"(row[4]), "v"(row[5]), "v"(row[6]),
          "v"(row[7]), "v"(row[8]), "v"(x), "i"(J));
  }
}

}  // namespace sdg
