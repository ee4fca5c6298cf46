// HIP backend of the gridding benchmark: the hardware_type::gpu specialisations of
// gridding_benchmark_launcher / degridding_benchmark_launcher, implemented over libidgx's C ABI (include/idgx.h).
//
// Replaces gridding/src/gridding_benchmark.cu:248-361 and gridding/src/degridding_benchmark.cu:205-318.  The file
// uses nothing but the reference's backend interface (configuration fields, Array*::data()/size(), Metadata,
// benchmark_result), so it compiles unchanged against the reference's own headers: a maintainer swaps the two
// .cu files for this one.  Per launch it allocates the device buffers, times `niterations` host-to-device uploads,
// `niterations` kernel runs and `niterations` copies of the result back with HIP events, frees the buffers, and
// reports per-iteration microseconds and MB/s.  Plain FP32 runs the one-shot kernels; the mixed forms run as a
// one-class batch (idgx_batch_*), created before the timed kernel runs.  Unlike the CUDA backend it copies the
// actual outputs back (subgrids after gridding, visibilities after degridding) and does not divide the
// bandwidth by the iteration count twice.  Errors print the runtime's message and exit with its status, as the
// CUDA backend's cudaCheck does (gridding/inc/util.cuh:4-17).
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>

#include "degridding_benchmark.h"
#include "gridding_benchmark.h"
#include "idgx.h"

namespace gridding_benchmark
{
    namespace
    {
        using benchmarks_common::hardware_type;

        void hip_check(hipError_t e, const char *what, int line)
        {
            if (e != hipSuccess)
            {
                std::fprintf(stderr, "GPUassert: %s %s %s %d\n", hipGetErrorString(e), what, __FILE__, line);
                std::exit(static_cast<int>(e));
            }
        }

        void idgx_check(int status, const char *what, int line)
        {
            if (status != IDGX_OK)
            {
                std::fprintf(stderr, "GPUassert: %s %s %s %d\n", idgx_strerror(status), what, __FILE__, line);
                std::exit(status);
            }
        }

#define HIP_CHECK(x) hip_check((x), #x, __LINE__)
#define IDGX_CHECK(x) idgx_check((x), #x, __LINE__)

        // Optional configuration keys of this backend, read when the configuration type has them (this
        // repository's driver) and otherwise taken from the environment / the reference's hard-coded values
        // (gridding/src/gridding_benchmark.cu:315-316).
        template <class C, class = void> struct has_backend_keys : std::false_type
        {
        };
        template <class C>
        struct has_backend_keys<C, std::void_t<decltype(std::declval<C>().precision),
                                               decltype(std::declval<C>().image_size),
                                               decltype(std::declval<C>().w_step_in_lambda)>> : std::true_type
        {
        };

        int precision_code(const std::string &p)
        {
            if (p == "fp32")
                return IDGX_PLAIN_FP32;
            if (p == "bf16x3")
                return IDGX_MIXED_BF16X3;
            if (p == "fp16x3")
                return IDGX_MIXED_FP16X3;
            if (p == "bf16")
                return IDGX_MIXED_BF16;
            if (p == "fp16")
                return IDGX_MIXED_FP16;
            std::fprintf(stderr, "Error: unknown precision %s (fp32, bf16x3, fp16x3, bf16, fp16)\n", p.c_str());
            std::exit(-1);
        }

        template <class C> void backend_keys(const C &c, int &precision, float &image_size, float &w_step)
        {
            if constexpr (has_backend_keys<C>::value)
            {
                precision = precision_code(c.precision);
                image_size = c.image_size;
                w_step = c.w_step_in_lambda;
            }
            else
            {
                const char *env = std::getenv("IDGX_PRECISION");
                precision = precision_code(env ? env : "fp32");
                image_size = 0.1F;
                w_step = 0.F;
            }
        }

        struct DeviceBuffer
        {
            void *ptr = nullptr;
            size_t bytes = 0;
            explicit DeviceBuffer(size_t n) : bytes(n)
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

        struct BatchGuard
        {
            idgx_batch *b = nullptr;
            BatchGuard() = default;
            BatchGuard(const BatchGuard &) = delete;
            BatchGuard &operator=(const BatchGuard &) = delete;
            ~BatchGuard()
            {
                idgx_batch_destroy(b);
            }
        };

        template <bool Gridding, class Config, class UVW, class Vis, class Aterms, class Floats, class Sph, class Sub,
                  class Meta>
        benchmark_result run(const Config &configuration, UVW &uvw, Vis &visibilities, Aterms &aterms,
                             Floats &wavenumbers, Sph &spheroidal, Sub &subgrids, Meta &metadata)
        {
            const size_t niter = configuration.niterations > 0 ? configuration.niterations : 1;
            int precision;
            float image_size, w_step;
            backend_keys(configuration, precision, image_size, w_step);

            const int S = static_cast<int>(configuration.subgrid_size);
            const size_t n_subgrids = metadata.size();
            const size_t n_stations = configuration.nstations;
            const idgx_metadata *h_md = reinterpret_cast<const idgx_metadata *>(metadata.data());
            int max_t = 0;
            for (size_t s = 0; s < n_subgrids; ++s)
                max_t = h_md[s].nr_timesteps > max_t ? h_md[s].nr_timesteps : max_t;

            idgx_plan plan{};
            plan.nr_subgrids = static_cast<int32_t>(n_subgrids);
            plan.subgrid_size = S;
            plan.grid_size = static_cast<int32_t>(configuration.grid_size);
            plan.nr_channels = static_cast<int32_t>(wavenumbers.size());
            plan.nr_stations = static_cast<int32_t>(n_stations);
            plan.nr_aterm_slots = static_cast<int32_t>(aterms.size() / (n_stations * S * S));
            plan.nr_rows = static_cast<int64_t>(uvw.size());
            plan.max_nr_timesteps = max_t;
            plan.image_size = image_size;
            plan.w_step_in_lambda = w_step;
            plan.precision = precision;
            plan.aterm_convention = IDGX_ATERM_ADJOINT; // as the CUDA degridder (common_gridding.cuh:82-94)
            plan.flags = 0;
            IDGX_CHECK(idgx_check_plan(&plan, h_md, nullptr));

            const size_t b_uvw = uvw.size() * sizeof(uvw[0]);
            const size_t b_wn = wavenumbers.size() * sizeof(wavenumbers[0]);
            const size_t b_vis = visibilities.size() * sizeof(visibilities[0]);
            const size_t b_sph = spheroidal.size() * sizeof(spheroidal[0]);
            const size_t b_at = aterms.size() * sizeof(aterms[0]);
            const size_t b_md = metadata.size() * sizeof(metadata[0]);
            const size_t b_sub = subgrids.size() * sizeof(subgrids[0]);
            DeviceBuffer d_uvw(b_uvw), d_wn(b_wn), d_vis(b_vis), d_sph(b_sph), d_at(b_at), d_md(b_md), d_sub(b_sub);

            // Inputs: the gridder reads visibilities, the degridder reads subgrids.
            const size_t in_bytes = b_uvw + b_wn + b_sph + b_at + b_md + (Gridding ? b_vis : b_sub);
            const size_t out_bytes = Gridding ? b_sub : b_vis;
            if (!Gridding)
                HIP_CHECK(hipMemset(d_vis.ptr, 0, b_vis));

            // The mixed forms run as a batch (persistent work-queue kernels, tile shapes chosen per subgrid size,
            // degridder pixels prepared once per subgrid); creating it builds the queue and allocates its device
            // memory, outside the timed region.  Plain FP32 runs the one-shot kernels.
            BatchGuard batch;
            idgx_batch_data bd{};
            if (precision != IDGX_PLAIN_FP32)
            {
                idgx_batch_class cls{};
                cls.plan = plan;
                cls.h_metadata = h_md;
                IDGX_CHECK(idgx_batch_create(&cls, 1, &batch.b));
                bd.d_uvw = static_cast<const float *>(d_uvw.ptr);
                bd.d_wavenumbers = static_cast<const float *>(d_wn.ptr);
                bd.d_visibilities = d_vis.ptr;
                bd.d_visibilities_out = d_vis.ptr;
                bd.cls[0].d_metadata = static_cast<const idgx_metadata *>(d_md.ptr);
                bd.cls[0].d_spheroidal = static_cast<const float *>(d_sph.ptr);
                bd.cls[0].d_aterms = d_at.ptr;
                bd.cls[0].d_subgrids = d_sub.ptr;
            }

            benchmark_result result;
            Events in, compute, out;
            HIP_CHECK(hipEventRecord(in.begin, nullptr));
            for (size_t i = 0; i < niter; ++i)
            {
                HIP_CHECK(hipMemcpy(d_uvw.ptr, uvw.data(), b_uvw, hipMemcpyHostToDevice));
                HIP_CHECK(hipMemcpy(d_wn.ptr, wavenumbers.data(), b_wn, hipMemcpyHostToDevice));
                HIP_CHECK(hipMemcpy(d_sph.ptr, spheroidal.data(), b_sph, hipMemcpyHostToDevice));
                HIP_CHECK(hipMemcpy(d_at.ptr, aterms.data(), b_at, hipMemcpyHostToDevice));
                HIP_CHECK(hipMemcpy(d_md.ptr, metadata.data(), b_md, hipMemcpyHostToDevice));
                if (Gridding)
                    HIP_CHECK(hipMemcpy(d_vis.ptr, visibilities.data(), b_vis, hipMemcpyHostToDevice));
                else
                    HIP_CHECK(hipMemcpy(d_sub.ptr, subgrids.data(), b_sub, hipMemcpyHostToDevice));
            }
            HIP_CHECK(hipEventRecord(in.end, nullptr));
            const double in_us = in.elapsed_us();

            HIP_CHECK(hipEventRecord(compute.begin, nullptr));
            for (size_t i = 0; i < niter; ++i)
            {
                if (batch.b)
                    IDGX_CHECK(Gridding ? idgx_batch_grid(batch.b, &bd, nullptr) : idgx_batch_degrid(batch.b, &bd, nullptr));
                else if (Gridding)
                    IDGX_CHECK(idgx_grid(&plan, static_cast<const idgx_metadata *>(d_md.ptr), nullptr,
                                         static_cast<const float *>(d_uvw.ptr), static_cast<const float *>(d_wn.ptr),
                                         d_vis.ptr, static_cast<const float *>(d_sph.ptr), d_at.ptr, d_sub.ptr,
                                         nullptr));
                else
                    IDGX_CHECK(idgx_degrid(&plan, static_cast<const idgx_metadata *>(d_md.ptr), nullptr,
                                           static_cast<const float *>(d_uvw.ptr), static_cast<const float *>(d_wn.ptr),
                                           d_sub.ptr, static_cast<const float *>(d_sph.ptr), d_at.ptr, d_vis.ptr,
                                           nullptr));
            }
            HIP_CHECK(hipEventRecord(compute.end, nullptr));
            const double compute_us = compute.elapsed_us();
            HIP_CHECK(hipGetLastError());

            HIP_CHECK(hipEventRecord(out.begin, nullptr));
            for (size_t i = 0; i < niter; ++i)
            {
                if (Gridding)
                    HIP_CHECK(hipMemcpy(subgrids.data(), d_sub.ptr, b_sub, hipMemcpyDeviceToHost));
                else
                    HIP_CHECK(hipMemcpy(visibilities.data(), d_vis.ptr, b_vis, hipMemcpyDeviceToHost));
            }
            HIP_CHECK(hipEventRecord(out.end, nullptr));
            const double out_us = out.elapsed_us();

            result.in_transfer_time = in_us / niter;
            result.out_transfer_time = out_us / niter;
            result.compute_time = compute_us / niter;
            // bytes per microsecond = MB/s
            result.in_bandwidth = in_us > 0 ? static_cast<double>(in_bytes) * niter / in_us : -1.;
            result.out_bandwidth = out_us > 0 ? static_cast<double>(out_bytes) * niter / out_us : -1.;
            return result;
        }
    } // namespace

    template <>
    benchmark_result gridding_benchmark_launcher<benchmarks_common::hardware_type::gpu>::launch(
        const configuration &configuration, Array2D<UVWCoordinate<float>> &uvw,
        Array3D<Visibility<std::complex<float>>> &visibilities, Array1D<Baseline> &baselines,
        Array4D<Matrix2x2<std::complex<float>>> &aterms, Array1D<float> &frequencies, Array1D<float> &wavenumbers,
        Array2D<float> &spheroidal, Array4D<std::complex<float>> &subgrids, Array1D<Metadata> &metadata)
    {
        (void)baselines;
        (void)frequencies;
        return run<true>(configuration, uvw, visibilities, aterms, wavenumbers, spheroidal, subgrids, metadata);
    }

    template <>
    benchmark_result degridding_benchmark_launcher<benchmarks_common::hardware_type::gpu>::launch(
        const configuration &configuration, Array2D<UVWCoordinate<float>> &uvw,
        Array3D<Visibility<std::complex<float>>> &visibilities, Array1D<Baseline> &baselines,
        Array4D<Matrix2x2<std::complex<float>>> &aterms, Array1D<float> &frequencies, Array1D<float> &wavenumbers,
        Array2D<float> &spheroidal, Array4D<std::complex<float>> &subgrids, Array1D<Metadata> &metadata)
    {
        (void)baselines;
        (void)frequencies;
        return run<false>(configuration, uvw, visibilities, aterms, wavenumbers, spheroidal, subgrids, metadata);
    }
} // namespace gridding_benchmark
