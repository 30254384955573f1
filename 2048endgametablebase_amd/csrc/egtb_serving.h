// Table sets resident in HBM (serving / reader) and the lookup launcher.
#pragma once

#include <memory>
#include <vector>

#include "egtb.h"
#include "egtb_ai.h"
#include "egtb_kernels.h"
#include "egtb_runtime.h"

namespace egtb {

struct TableSet {
    int dev = 0;
    int steps = 0;
    PatternDev pattern{};
    int64_t nums_adjust = 0;  // minus the seed tile sum
    int rate_size = 4;
    bool is_float = true;
    double scale = 1.0;  // raw rate / scale = probability
    double zero = 0.0;   // the value of a key absent from its layer (-1 for the 1-float types)
    double p4 = 0.1;
    // evaluate every symmetry op and keep the one whose best rate is largest
    // (> 0); reference prefer_max_result_ (ReaderRuntime.cpp:827-832,1053)
    int prefer_max = 0;
    std::vector<uint64_t> n;  // records per layer
    Arena storage{(size_t)256 << 20};
    DBuf tabs;  // TableDev<T>[steps] on the device
    size_t bytes = 0;
    const Result *borrowed = nullptr;  // tables owned by a build result
    AiTables ai{};                     // the off-table fallback's evaluation tables on this device
};

void configure_serving_kernels();
template <typename OutT>
void launch_lookup(hipStream_t s, const TableSet &ts, const uint64_t *boards, uint64_t n, OutT *out4,
                   int32_t *status, int32_t *op_out);

std::unique_ptr<TableSet> table_from_files(const char *prefix, int steps, const egtb_pattern *pat, int target,
                                           int64_t tile_sum, const char *dtype, double p4, int device);
std::unique_ptr<TableSet> table_from_result(const Result &r, const egtb_pattern *pat, int target,
                                            int64_t tile_sum, double p4);

}  // namespace egtb
