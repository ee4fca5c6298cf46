// Host data types of the gridding benchmark, layout-compatible with the reference's gridding/inc/types.hpp:7-56
// (plan record, 2x2 correlation matrices, uvw) and with the same owning N-d array interface (data(), size(),
// get_*_dim(), operator() row-major indexing), so backend translation units compile unchanged against either.
#pragma once
#include <complex>
#include <cstddef>
#include <cstdlib>
#include <new>
#include <utility>

namespace gridding_benchmark
{
    struct Coordinate
    {
        int x, y, z;
    };

    struct Baseline
    {
        unsigned int station1, station2;
    };

    struct Metadata
    {
        int baseline_offset;
        int time_offset;
        int nr_timesteps;
        int aterm_index;
        Baseline baseline;
        Coordinate coordinate;
    };
    static_assert(sizeof(Metadata) == 36, "Metadata must keep the reference's 36-byte layout");

    template <class T> struct Matrix2x2
    {
        T xx, xy, yx, yy;
    };

    template <class T> using Visibility = Matrix2x2<T>;

    template <class T> struct UVWCoordinate
    {
        T u, v, w;
    };

    namespace detail
    {
        // 64-byte aligned, zero-initialised owning buffer.
        template <class T> class Buffer
        {
          public:
            explicit Buffer(size_t n) : m_n(n), m_ptr(nullptr)
            {
                if (n == 0)
                    return;
                const size_t bytes = ((n * sizeof(T) + 63) / 64) * 64;
                m_ptr = static_cast<T *>(std::aligned_alloc(64, bytes));
                if (!m_ptr)
                    throw std::bad_alloc();
                for (size_t i = 0; i < n; ++i)
                    new (m_ptr + i) T();
            }
            Buffer(const Buffer &) = delete;
            Buffer &operator=(const Buffer &) = delete;
            Buffer(Buffer &&o) noexcept : m_n(o.m_n), m_ptr(o.m_ptr)
            {
                o.m_ptr = nullptr;
                o.m_n = 0;
            }
            Buffer &operator=(Buffer &&o) noexcept
            {
                std::swap(m_n, o.m_n);
                std::swap(m_ptr, o.m_ptr);
                return *this;
            }
            ~Buffer()
            {
                std::free(m_ptr);
            }
            T *get() const
            {
                return m_ptr;
            }
            size_t count() const
            {
                return m_n;
            }

          private:
            size_t m_n;
            T *m_ptr;
        };
    } // namespace detail

    template <class T> class Array1D
    {
      public:
        explicit Array1D(size_t x = 0) : m_x(x), m_buf(x)
        {
        }
        T *data(size_t i = 0) const
        {
            return m_buf.get() + i;
        }
        size_t size() const
        {
            return m_x;
        }
        size_t bytes() const
        {
            return m_x * sizeof(T);
        }
        size_t get_x_dim() const
        {
            return m_x;
        }
        T &operator()(size_t i)
        {
            return data()[i];
        }
        const T &operator()(size_t i) const
        {
            return data()[i];
        }
        T &operator[](size_t i)
        {
            return data()[i];
        }
        const T &operator[](size_t i) const
        {
            return data()[i];
        }

      private:
        size_t m_x;
        detail::Buffer<T> m_buf;
    };

    template <class T> class Array2D
    {
      public:
        Array2D(size_t y = 0, size_t x = 0) : m_y(y), m_x(x), m_buf(y * x)
        {
        }
        T *data(size_t i = 0) const
        {
            return m_buf.get() + i;
        }
        size_t size() const
        {
            return m_y * m_x;
        }
        size_t bytes() const
        {
            return size() * sizeof(T);
        }
        size_t get_x_dim() const
        {
            return m_x;
        }
        size_t get_y_dim() const
        {
            return m_y;
        }
        T &operator()(size_t y, size_t x)
        {
            return data()[y * m_x + x];
        }
        const T &operator()(size_t y, size_t x) const
        {
            return data()[y * m_x + x];
        }
        T &operator[](size_t i)
        {
            return data()[i];
        }
        const T &operator[](size_t i) const
        {
            return data()[i];
        }

      private:
        size_t m_y, m_x;
        detail::Buffer<T> m_buf;
    };

    template <class T> class Array3D
    {
      public:
        Array3D(size_t z = 0, size_t y = 0, size_t x = 0) : m_z(z), m_y(y), m_x(x), m_buf(z * y * x)
        {
        }
        T *data(size_t i = 0) const
        {
            return m_buf.get() + i;
        }
        size_t size() const
        {
            return m_z * m_y * m_x;
        }
        size_t bytes() const
        {
            return size() * sizeof(T);
        }
        size_t get_x_dim() const
        {
            return m_x;
        }
        size_t get_y_dim() const
        {
            return m_y;
        }
        size_t get_z_dim() const
        {
            return m_z;
        }
        T &operator()(size_t z, size_t y, size_t x)
        {
            return data()[(z * m_y + y) * m_x + x];
        }
        const T &operator()(size_t z, size_t y, size_t x) const
        {
            return data()[(z * m_y + y) * m_x + x];
        }
        T &operator[](size_t i)
        {
            return data()[i];
        }
        const T &operator[](size_t i) const
        {
            return data()[i];
        }

      private:
        size_t m_z, m_y, m_x;
        detail::Buffer<T> m_buf;
    };

    template <class T> class Array4D
    {
      public:
        Array4D(size_t w = 0, size_t z = 0, size_t y = 0, size_t x = 0)
            : m_w(w), m_z(z), m_y(y), m_x(x), m_buf(w * z * y * x)
        {
        }
        T *data(size_t i = 0) const
        {
            return m_buf.get() + i;
        }
        size_t size() const
        {
            return m_w * m_z * m_y * m_x;
        }
        size_t bytes() const
        {
            return size() * sizeof(T);
        }
        size_t get_x_dim() const
        {
            return m_x;
        }
        size_t get_y_dim() const
        {
            return m_y;
        }
        size_t get_z_dim() const
        {
            return m_z;
        }
        size_t get_w_dim() const
        {
            return m_w;
        }
        T &operator()(size_t w, size_t z, size_t y, size_t x)
        {
            return data()[((w * m_z + z) * m_y + y) * m_x + x];
        }
        const T &operator()(size_t w, size_t z, size_t y, size_t x) const
        {
            return data()[((w * m_z + z) * m_y + y) * m_x + x];
        }
        T &operator[](size_t i)
        {
            return data()[i];
        }
        const T &operator[](size_t i) const
        {
            return data()[i];
        }

      private:
        size_t m_w, m_z, m_y, m_x;
        detail::Buffer<T> m_buf;
    };
} // namespace gridding_benchmark
