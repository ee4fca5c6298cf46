// HIP backend of the FFT benchmark: the GPU launch of fft/src/fft_benchmark.cpp:240-345 over libidgx's idgx_fft2d
// (include/idgx.h): this library's kernels for every plane with power-of-two sides in 16 .. 512 (FP32 and FP64; one
// pass up to 128 x 128, in registers for FP32 64 x 64 and 128 x 128), FP32 planes with power-of-two sides in 1024 ..
// 16384 (three passes), FP32 square planes of side 3 x 2^k, 3072 .. 24576 (three passes, each column pass opening with
// a radix-3 step) and the other power-of-two shapes with rows of up to 16384 (FP32) / 8192 (FP64) points (one-pass
// rows, four-step columns); rocFFT for the rest (idgx_fft2d_own_kernels tells which).  The sequence is the reference's:
// as many planes as fit in memorysize (input + output planes; the reference also counts heFFTe's workspace), a first
// transform of the batch timed as initialisation (plan or scratch creation; heFFTe's plan does not depend on the batch,
// rocFFT's does), niterations uploads of the batch, niterations transforms of the batch, niterations downloads, then
// the inverse transform of the output and the maximum error of the first plane against its input, the pattern k mod 128
// (fft/src/fft_benchmark.cpp:82-108).  Times are per transform (total / (batch x niterations)), bandwidths in bytes per
// microsecond.  HIP events time the device work.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "fft_configuration.h"
#include "idgx.h"

namespace fft_benchmark
{
    namespace
    {
        void hip_check(hipError_t e, const char *what, int line)
        {
            if (e != hipSuccess)
            {
                std::fprintf(stderr, "GPUassert: %s %s %s %d\n", hipGetErrorString(e), what, __FILE__, line);
                std::exit(static_cast<int>(e));
            }
        }
#define HIP_CHECK(x) hip_check((x), #x, __LINE__)

        void idgx_check(int status, const char *what, int line)
        {
            if (status != IDGX_OK)
            {
                std::fprintf(stderr, "GPUassert: %s %s %s %d\n", idgx_strerror(status), what, __FILE__, line);
                std::exit(status);
            }
        }
#define IDGX_CHECK(x) idgx_check((x), #x, __LINE__)

        struct DeviceBuffer
        {
            void *ptr = nullptr;
            explicit DeviceBuffer(size_t n)
            {
                HIP_CHECK(hipMalloc(&ptr, n > 0 ? n : 1));
            }
            DeviceBuffer(const DeviceBuffer &) = delete;
            DeviceBuffer &operator=(const DeviceBuffer &) = delete;
            ~DeviceBuffer()
            {
                if (ptr)
                    (void)hipFree(ptr);
            }
        };

        struct Events
        {
            hipEvent_t begin, end;
            Events()
            {
                HIP_CHECK(hipEventCreate(&begin));
                HIP_CHECK(hipEventCreate(&end));
            }
            ~Events()
            {
                (void)hipEventDestroy(begin);
                (void)hipEventDestroy(end);
            }
            double elapsed_us()
            {
                HIP_CHECK(hipEventSynchronize(end));
                float ms = 0.F;
                HIP_CHECK(hipEventElapsedTime(&ms, begin, end));
                return 1000.0 * ms;
            }
        };

        template <typename T> benchmark_result run(const configuration &c)
        {
            using complex = std::complex<T>;
            const size_t plane = c.nx * c.ny;
            const size_t plane_bytes = plane * sizeof(complex);
            if (c.nx == 0 || c.ny == 0 || c.nx > INT_MAX || c.ny > INT_MAX || plane > INT_MAX)
                return benchmark_result{};
            const size_t batch = c.memorysize / (2 * plane_bytes);
            if (batch == 0)
                return benchmark_result{};
            const int nx = static_cast<int>(c.nx), ny = static_cast<int>(c.ny);
            const int precision = sizeof(T) == sizeof(double) ? IDGX_FFT_FP64 : IDGX_FFT_FP32;
            const int sign = static_cast<int>(c.ttype);
            const size_t niter = std::max<size_t>(c.niterations, 1);

            DeviceBuffer d_in(batch * plane_bytes), d_out(batch * plane_bytes);
            HIP_CHECK(hipMemset(d_in.ptr, 0, batch * plane_bytes));
            HIP_CHECK(hipDeviceSynchronize());
            // Initialisation: one transform of the whole batch, which creates what the transform keeps (the rocFFT
            // plan for this shape and batch, or the three-pass kernels' scratch).
            const auto i0 = std::chrono::steady_clock::now();
            IDGX_CHECK(idgx_fft2d(d_in.ptr, d_out.ptr, nx, ny, static_cast<int64_t>(batch), sign, precision, nullptr));
            HIP_CHECK(hipDeviceSynchronize());
            const double init_us =
                std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - i0).count();

            std::vector<complex> in(batch * plane);
            for (size_t k = 0; k < plane; ++k)
                in[k] = complex(static_cast<T>(k % 128));

            Events up, compute, down;
            HIP_CHECK(hipEventRecord(up.begin, nullptr));
            for (size_t i = 0; i < niter; ++i)
                HIP_CHECK(hipMemcpy(d_in.ptr, in.data(), batch * plane_bytes, hipMemcpyHostToDevice));
            HIP_CHECK(hipEventRecord(up.end, nullptr));
            const double in_us = up.elapsed_us();

            HIP_CHECK(hipEventRecord(compute.begin, nullptr));
            for (size_t i = 0; i < niter; ++i)
                IDGX_CHECK(idgx_fft2d(d_in.ptr, d_out.ptr, nx, ny, static_cast<int64_t>(batch), sign, precision,
                                      nullptr));
            HIP_CHECK(hipEventRecord(compute.end, nullptr));
            const double compute_us = compute.elapsed_us();

            std::vector<complex> out(batch * plane);
            HIP_CHECK(hipEventRecord(down.begin, nullptr));
            for (size_t i = 0; i < niter; ++i)
                HIP_CHECK(hipMemcpy(out.data(), d_out.ptr, batch * plane_bytes, hipMemcpyDeviceToHost));
            HIP_CHECK(hipEventRecord(down.end, nullptr));
            const double out_us = down.elapsed_us();

            // Validation: the inverse transform of the output back into the input buffer, first plane.
            IDGX_CHECK(idgx_fft2d(d_out.ptr, d_in.ptr, nx, ny, static_cast<int64_t>(batch), -sign, precision,
                                  nullptr));
            HIP_CHECK(hipMemcpy(in.data(), d_in.ptr, plane_bytes, hipMemcpyDeviceToHost));
            const T scale = T(1) / static_cast<T>(plane);
            T max_error = 0;
            for (size_t k = 0; k < plane; ++k)
                max_error = std::max(max_error, std::abs(complex(static_cast<T>(k % 128)) - in[k] * scale));

            const double n = static_cast<double>(batch * niter);
            benchmark_result r;
            r.status = benchmark_result::status_t::success;
            r.batch_size = batch;
            r.niterations = niter;
            r.init_time = init_us;
            r.in_transfer_time = in_us / n;
            r.out_transfer_time = out_us / n;
            r.compute_time = compute_us / n;
            r.in_bandwidth = in_us > 0 ? static_cast<double>(batch * plane_bytes) * niter / in_us : -1.;
            r.out_bandwidth = out_us > 0 ? static_cast<double>(batch * plane_bytes) * niter / out_us : -1.;
            r.max_error = static_cast<double>(max_error);
            return r;
        }
    } // namespace

    benchmark_result launch_benchmark(const configuration &c)
    {
        if (c.htype != benchmarks_common::hardware_type::gpu)
            return benchmark_result{};
        // double runs FP64 planes; the reference's GPU branch for double instantiates its single-precision
        // transform instead (fft/src/fft_benchmark.cpp:62-65).
        return c.ftype == float_type::double_precision ? run<double>(c) : run<float>(c);
    }
} // namespace fft_benchmark
