// Error handling, version query and the line catalogue of librbx.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "catalog.hpp"

namespace rbx {

static thread_local std::string g_error;

void set_error(const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
}

const std::string &last_error() { return g_error; }

int read_launch(const rbx_launch *in, rbx_launch &out, const char *who) {
    out = rbx_launch{};
    if (in == nullptr) return RBX_OK;
    out = *in;
    RBX_CHECK_ARG(out.alpha_k == 0 || out.alpha_k == 1 || out.alpha_k == 2 || out.alpha_k == 4,
                  "%s: launch.alpha_k=%d (0, 1, 2 or 4)", who, out.alpha_k);
    RBX_CHECK_ARG(out.alpha_series >= 0 && out.alpha_series <= 2, "%s: launch.alpha_series=%d (0, 1 or 2)", who,
                  out.alpha_series);
    RBX_CHECK_ARG(out.alpha_split >= 0 && out.alpha_split <= kMaxLaunchSplit, "%s: launch.alpha_split=%d (0..%d)",
                  who, out.alpha_split, kMaxLaunchSplit);
    RBX_CHECK_ARG(out.far_split >= 0 && out.far_split <= kMaxLaunchSplit, "%s: launch.far_split=%d (0..%d)", who,
                  out.far_split, kMaxLaunchSplit);
    RBX_CHECK_ARG(out.int_unroll == 0 || out.int_unroll == 4 || out.int_unroll == 8,
                  "%s: launch.int_unroll=%d (0, 4 or 8)", who, out.int_unroll);
    RBX_CHECK_ARG(out.int_map >= 0 && out.int_map <= 2, "%s: launch.int_map=%d (0, 1 or 2)", who, out.int_map);
    RBX_CHECK_ARG(out.tan_split >= 0 && out.tan_split <= kMaxLaunchSplit, "%s: launch.tan_split=%d (0..%d)", who,
                  out.tan_split, kMaxLaunchSplit);
    RBX_CHECK_ARG(out.adj_rays == 0 || out.adj_rays == 1 || out.adj_rays == 2 || out.adj_rays == 4 ||
                      out.adj_rays == 8,
                  "%s: launch.adj_rays=%d (0, 1, 2, 4 or 8)", who, out.adj_rays);
    RBX_CHECK_ARG(out.adj_copies >= 0 && out.adj_copies <= 64, "%s: launch.adj_copies=%d (0..64)", who,
                  out.adj_copies);
    for (int r : out.reserved) RBX_CHECK_ARG(r == 0, "%s: launch.reserved must be 0", who);
    return RBX_OK;
}

}  // namespace rbx

using rbx::kNumCols;

extern "C" int rbx_abi_version(void) { return RBX_ABI_VERSION; }

extern "C" int rbx_last_error(char *buf, size_t len) {
    if (buf == nullptr || len == 0) return RBX_ERR_ARG;
    const std::string &e = rbx::last_error();
    size_t n = e.size() < len - 1 ? e.size() : len - 1;
    std::memcpy(buf, e.data(), n);
    buf[n] = '\0';
    return RBX_OK;
}

extern "C" int rbx_device_count(int *count) {
    RBX_CHECK_ARG(count != nullptr, "rbx_device_count: null count");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    *count = (e == hipSuccess) ? n : 0;
    return RBX_OK;
}

extern "C" int rbx_catalog_create(rbx_catalog **out) {
    RBX_CHECK_ARG(out != nullptr, "rbx_catalog_create: null out");
    auto *c = new (std::nothrow) rbx_catalog;
    if (c == nullptr) {
        rbx::set_error("rbx_catalog_create: host allocation failed");
        return RBX_ERR_NOMEM;
    }
    std::memset(c, 0, sizeof(*c));
    *out = c;
    return RBX_OK;
}

static void free_block(rbx_catalog *cat, int kind) {
    if (cat->storage[kind] != nullptr) {
        (void)hipFree(cat->storage[kind]);
        cat->storage[kind] = nullptr;
    }
    std::memset(&cat->dev.blk[kind], 0, sizeof(rbx::DevBlock));
}

// Strength factor bounding a block's line weights (see DevBlock::wscale).
static double block_wscale(int kind, int n, const double *const *col) {
    double m = 0.0;
    for (int i = 0; i < n; ++i) {
        double f0 = col[0][i];
        double s;
        switch (kind) {
            case RBX_LINES_NH3_ROT:
            case RBX_LINES_NH3_V2: s = col[1][i]; break;
            case RBX_LINES_PH3: s = col[1][i] * col[3][i] / (f0 * f0); break;
            default: s = col[1][i] / (f0 * f0); break;
        }
        s = std::fabs(s);
        if (std::isfinite(s) && s > m) m = s;
    }
    return m;
}

extern "C" int rbx_catalog_set_lines(rbx_catalog *cat, int kind, int n, const double *const *host_columns) {
    RBX_CHECK_ARG(cat != nullptr, "rbx_catalog_set_lines: null catalogue");
    RBX_CHECK_ARG(kind >= 0 && kind < RBX_NUM_LINE_KINDS, "rbx_catalog_set_lines: bad kind %d", kind);
    RBX_CHECK_ARG(n >= 0, "rbx_catalog_set_lines: negative line count %d", n);
    RBX_CHECK_ARG(n == 0 || host_columns != nullptr, "rbx_catalog_set_lines: null columns");
    const int ncol = kNumCols[kind];
    for (int i = 0; n > 0 && i < ncol; ++i)
        RBX_CHECK_ARG(host_columns[i] != nullptr, "rbx_catalog_set_lines: column %d is null", i);
    for (int i = 0; i < n; ++i)
        RBX_CHECK_ARG(host_columns[0][i] > 0.0, "rbx_catalog_set_lines: line %d has centre %g <= 0",
                      i, host_columns[0][i]);
    if (kind == RBX_TABLE_H2_ORTON && n > 0) {
        // the kernels bracket a frequency by table rows i-1, i, i+1 and a temperature by a
        // table interval: both grids must be ascending
        RBX_CHECK_ARG(n >= 3, "rbx_catalog_set_lines: the Orton table needs >= 3 frequencies, got %d", n);
        for (int i = 1; i < n; ++i)
            RBX_CHECK_ARG(host_columns[0][i] > host_columns[0][i - 1],
                          "rbx_catalog_set_lines: Orton table frequencies not ascending at %d", i);
        for (int i = 1; i < RBX_ORTON_NTEMP; ++i)
            RBX_CHECK_ARG(host_columns[2][i] > host_columns[2][i - 1] && host_columns[2][0] > 0.0,
                          "rbx_catalog_set_lines: Orton table temperatures not positive and ascending");
    }
    free_block(cat, kind);
    if (n == 0) return RBX_OK;
    size_t total = 0;
    for (int c = 0; c < ncol; ++c) total += rbx::column_length(kind, c, n);
    const size_t bytes = sizeof(double) * total;
    std::vector<double> packed(total);
    // the line lists summed by the forward line-sum kernel are stored by ascending
    // centre frequency (stable): neighbouring staged lines then sit at similar
    // distances from a wavefront's columns, which its FP32 fast path relies on
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    if (rbx::sorted_by_centre(kind))
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return host_columns[0][a] < host_columns[0][b]; });
    size_t at = 0;
    for (int c = 0; c < ncol; ++c) {
        const size_t len = rbx::column_length(kind, c, n);
        if (rbx::sorted_by_centre(kind)) {
            for (int i = 0; i < n; ++i) packed[at + i] = host_columns[c][order[i]];
        } else {
            std::memcpy(packed.data() + at, host_columns[c], sizeof(double) * len);
        }
        at += len;
    }
    double *d = nullptr;
    hipError_t e = hipMalloc(&d, bytes);
    if (e != hipSuccess) {
        rbx::set_error("rbx_catalog_set_lines: hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        return RBX_ERR_NOMEM;
    }
    e = hipMemcpy(d, packed.data(), bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        rbx::set_error("rbx_catalog_set_lines: upload failed: %s", hipGetErrorString(e));
        return RBX_ERR_HIP;
    }
    cat->storage[kind] = d;
    rbx::DevBlock &b = cat->dev.blk[kind];
    at = 0;
    for (int c = 0; c < ncol; ++c) {
        b.col[c] = d + at;
        at += rbx::column_length(kind, c, n);
    }
    b.n = n;
    b.wscale = kind == RBX_TABLE_H2_ORTON ? 0.0 : block_wscale(kind, n, host_columns);
    return RBX_OK;
}

extern "C" int rbx_catalog_lines(const rbx_catalog *cat, int kind, int *n) {
    RBX_CHECK_ARG(cat != nullptr && n != nullptr, "rbx_catalog_lines: null argument");
    RBX_CHECK_ARG(kind >= 0 && kind < RBX_NUM_LINE_KINDS, "rbx_catalog_lines: bad kind %d", kind);
    *n = cat->dev.blk[kind].n;
    return RBX_OK;
}

extern "C" int rbx_catalog_destroy(rbx_catalog *cat) {
    if (cat == nullptr) return RBX_OK;
    for (int k = 0; k < RBX_NUM_LINE_KINDS; ++k) free_block(cat, k);
    delete cat;
    return RBX_OK;
}
