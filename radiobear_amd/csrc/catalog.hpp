// Device-resident line catalogue shared by the absorption kernels.
#pragma once

#include "common.hpp"

namespace rbx {

constexpr int kMaxCols = 10;
constexpr int kNumCols[RBX_NUM_LINE_KINDS] = {4, 6, 3, 4, 4, 6, 10, 3, 3};

// values in column c of a block of n lines (n for line lists; the Orton table's
// columns are its frequencies, its [6 n][RBX_ORTON_NTEMP] values and its temperatures)
inline size_t column_length(int kind, int c, int n) {
    if (kind != RBX_TABLE_H2_ORTON) return static_cast<size_t>(n);
    if (c == 1) return static_cast<size_t>(6) * n * RBX_ORTON_NTEMP;
    return c == 2 ? static_cast<size_t>(RBX_ORTON_NTEMP) : static_cast<size_t>(n);
}

// line lists the forward line-sum kernel stages (every column has one value per line):
// stored sorted by centre frequency, column 0
inline bool sorted_by_centre(int kind) {
    return kind == RBX_LINES_NH3_INV || kind == RBX_LINES_NH3_ROT || kind == RBX_LINES_NH3_V2 ||
           kind == RBX_LINES_H2S || kind == RBX_LINES_PH3 || kind == RBX_LINES_H2O;
}

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
