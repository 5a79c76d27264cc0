// Device-resident line catalogue shared by the absorption kernels.
#pragma once

#include "common.hpp"

namespace rbx {

constexpr int kMaxCols = 10;
constexpr int kNumCols[RBX_NUM_LINE_KINDS] = {4, 6, 3, 4, 4, 6, 10};

// One block of lines as the kernels see it: SoA float64 columns in HBM.
struct DevBlock {
    const double *col[kMaxCols];
    int n;
    // max over lines of the strength factor that multiplies the line weight
    // (I0/f0^2, or I0 for the Gross blocks); sets the FP32 power-of-two scaling
    double wscale;
};

struct DevCatalog {
    DevBlock blk[RBX_NUM_LINE_KINDS];
};

}  // namespace rbx

struct rbx_catalog {
    rbx::DevCatalog dev;
    double *storage[RBX_NUM_LINE_KINDS];
};
