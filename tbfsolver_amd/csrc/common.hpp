// common.hpp -- shared internals of libtbfgpu: status handling, grid geometry,
// device buffers and the device-side view of one multigrid level.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/tbfgpu.h"

namespace tbfg {

// ------------------------------------------------------------------ status

void set_error(const char* fmt, ...);
void clear_error();

#define TBFG_HIP(...)                                                                             \
    do {                                                                                          \
        hipError_t e_ = (__VA_ARGS__);                                                            \
        if (e_ != hipSuccess) {                                                                   \
            ::tbfg::set_error("%s failed: %s (%s:%d)", #__VA_ARGS__, hipGetErrorString(e_), __FILE__, \
                              __LINE__);                                                          \
            return TBFG_EDEVICE;                                                                  \
        }                                                                                         \
    } while (0)

#define TBFG_TRY(...)                   \
    do {                                \
        int rc_ = (__VA_ARGS__);        \
        if (rc_ != TBFG_OK) return rc_; \
    } while (0)

// ------------------------------------------------------------------ geometry

// One axis of one grid level: centre and face coordinates with ghost layers,
// following grid.f90:235-399 (uniform and tanh distributions evaluated with the
// level's own cell count, so coarse centres coincide with fine faces 2i-1 and
// coarse faces with fine faces 2i, grid.f90:490-544).
struct Axis {
    static constexpr int G = 3;  // ghost layers
    int n = 0;
    double L = 0;
    int stretch = 0;
    double gamma = 1.4;
    std::vector<double> cbuf, fbuf;  // c(i), i = 1-G..n+G ; f(i), i = -G..n+G
    double c(int i) const { return cbuf[i - (1 - G)]; }
    double f(int i) const { return fbuf[i + G]; }
    double dc(int i) const { return c(i) - c(i - 1); }
    double df(int i) const { return f(i) - f(i - 1); }
    void build(int n_, double L_, int stretch_, double gamma_);
};

double coord_centre(int i, int n, double L, int stretch, double gamma);
double coord_face(int i, int n, double L, int stretch, double gamma);

struct HostLevel {
    int n[3];
    Axis ax[3];
    size_t cells() const { return (size_t)n[0] * n[1] * n[2]; }
};

int validate_grid(const tbfg_grid_desc* g);
// Cache key of a grid on a device, built from the descriptor's values (not its
// bytes, whose padding a caller may leave indeterminate).
std::string grid_key(const tbfg_grid_desc* g, int dev);
HostLevel make_level(const tbfg_grid_desc* g, int lev);

// ------------------------------------------------------------ device memory

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes)
    {
        o.p = nullptr;
        o.bytes = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    int alloc(size_t b);
    void release();
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

int upload(DevBuf& d, const void* src, size_t bytes);

// Synchronous copies and fills of handle creation and result reads, on a
// library-internal non-blocking stream of the current device (never ordered
// with the legacy null stream, so they stay legal while another thread
// captures a HIP graph on its own stream).
int copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind);
int memset_sync(void* dst, int value, size_t bytes);

// Per-axis device tables of one level, stored in T.  For cell i (1-based):
//   ip[i-1] = 1/(dc(i+1) df(i)), im[i-1] = 1/(dc(i) df(i))       (pcg.f90:448-463)
// restriction to the next coarser level, per coarse index I (1-based):
//   rlo[I-1] = (c_f(2I) - c_c(I)) / (c_f(2I) - c_f(2I-1)), rhi = 1 - ... (exact form)
//                                                            (interpolation.f90:97-118)
// prolongation from the next coarser level, per fine index i (1-based), coarse ih = i/2:
//   plo[i-1] = (c_c(ih+1) - c_f(i)) / (c_c(ih+1) - c_c(ih)), phi = (c_f(i) - c_c(ih)) / (...)
//                                                            (interpolation.f90:148-174)
template <class T>
struct AxisDev {
    const T* ip;
    const T* im;
    const T* rlo;
    const T* rhi;
    const T* plo;
    const T* phi;
};

template <class T>
struct LevelDev {
    int nx, ny, nz;
    AxisDev<T> a[3];
};

// host-side owner of the tables of one level in precision T
template <class T>
struct LevelTables {
    DevBuf buf;
    LevelDev<T> view{};
    int build(const HostLevel& L, const HostLevel* coarser);
};

// ---------------------------------------------------------- launch helpers

inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

inline int next_pow2(int v)
{
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

}  // namespace tbfg
