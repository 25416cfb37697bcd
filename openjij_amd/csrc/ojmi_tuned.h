// ojmi_tuned.h -- launch shapes chosen by openjij_amd/tune.py (generated;
// measured on AMD Instinct MI355X).  Re-run the tuner on another GPU and rebuild.
#pragma once
#include <stdint.h>

namespace ojmi {
namespace tuned {

// CSR graphs of degree <= 16: waves per replica, by field type (0 FP64,
// 1 FP32, 2 INT32) and replica count (first entry with R <= r_max).
struct CsrShape {
  int ftype;
  int64_t r_max;
  int waves;
};
constexpr CsrShape kCsr[] = {
    {0, 362, 4},
    {0, 724, 4},
    {0, 1448, 4},
    {0, 2896, 4},
    {0, 1099511627776, 4},
    {1, 362, 4},
    {1, 724, 4},
    {1, 1448, 4},
    {1, 2896, 4},
    {1, 1099511627776, 4},
    {2, 362, 4},
    {2, 724, 4},
    {2, 1448, 4},
    {2, 2896, 4},
    {2, 1099511627776, 4},
};

// Dense classical FP32 / INT32 systems: replicas per workgroup (1: one wave
// per replica), column-owned fields (1) or rows staged in LDS (0), XCD
// pacing window, by field type, 64-spin slot count and replica count
// (first entry with matching type and slots and R <= r_max).
struct DenseShape {
  int ftype;
  int slots;
  int64_t r_max;
  int group;
  int colown;
  int xcd_window;
};
constexpr DenseShape kDense[] = {
    {1, 8, 128, 1, 1, 0},
    {1, 8, 724, 1, 1, 0},
    {1, 8, 1099511627776, 1, 1, 0},
    {1, 16, 128, 1, 1, 0},
    {1, 16, 724, 1, 1, 0},
    {1, 16, 1099511627776, 1, 1, 0},
    {1, 32, 128, 4, 1, 0},
    {1, 32, 724, 4, 1, 0},
    {1, 32, 1099511627776, 4, 1, 0},
    {1, 64, 128, 4, 1, 2},
    {1, 64, 724, 4, 1, 2},
    {1, 64, 1099511627776, 4, 1, 2},
    {2, 8, 128, 1, 1, 0},
    {2, 8, 724, 1, 1, 0},
    {2, 8, 1099511627776, 1, 1, 0},
    {2, 16, 128, 1, 1, 0},
    {2, 16, 724, 1, 1, 0},
    {2, 16, 1099511627776, 1, 1, 0},
    {2, 32, 128, 1, 1, 0},
    {2, 32, 724, 1, 1, 0},
    {2, 32, 1099511627776, 1, 1, 0},
    {2, 64, 128, 1, 1, 0},
    {2, 64, 724, 1, 1, 0},
    {2, 64, 1099511627776, 1, 1, 0},
};

// Batches of problems: default launch (0 fused largest first, 1 per slot
// class, 2 persistent per-XCD queues).
constexpr int kBatchMode = 0;

// Dense FP32 systems with replicas to fill the device: per 16-sweep
// segment the matrix-core sweep while the acceptance rate is at least
// accept_permille per mille, the column-owned kernel below (0: never the
// matrix-core sweep), by 64-spin slot count.
struct MfmaChoice {
  int slots;
  int accept_permille;
};
constexpr MfmaChoice kMfma[] = {
    {32, 66},
    {64, 87},
};

}  // namespace tuned
}  // namespace ojmi
