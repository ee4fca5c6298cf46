// Result of one benchmark launch (reference: gridding/inc/common_gridding.h:5-12): times in microseconds per
// iteration, bandwidths in MB/s, -1 when not applicable.
#pragma once

namespace gridding_benchmark
{
    struct benchmark_result
    {
        double in_transfer_time{-1.};
        double out_transfer_time{-1.};
        double in_bandwidth{-1.};
        double out_bandwidth{-1.};
        double compute_time{-1.};
    };
} // namespace gridding_benchmark
