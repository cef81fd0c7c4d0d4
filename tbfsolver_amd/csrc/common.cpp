// common.cpp -- status handling, grid geometry and per-level device tables.
#include "common.hpp"

#include <cmath>
#include <mutex>
#include <cstdarg>
#include <cstring>

namespace tbfg {

static thread_local std::string g_last_error;

void set_error(const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

void clear_error() { g_last_error.clear(); }

const char* last_error() { return g_last_error.c_str(); }

// grid.f90:235-330.  The evaluation order follows the reference expressions so
// that coordinates agree to the last bit with a Fortran build using the same libm.
double coord_centre(int i, int n, double L, int stretch, double gamma)
{
    if (!stretch) {
        const double delta = L / n;
        return 0.5 * delta + (i - 1) * delta;
    }
    return 0.5 * L * (1.0 + std::tanh(-gamma + 2.0 * gamma * (i - 0.5) / n) / std::tanh(gamma));
}

double coord_face(int i, int n, double L, int stretch, double gamma)
{
    if (!stretch) {
        const double delta = L / n;
        return i * delta;
    }
    return 0.5 * L * (1.0 + std::tanh(-gamma + 2.0 * gamma * i / n) / std::tanh(gamma));
}

void Axis::build(int n_, double L_, int stretch_, double gamma_)
{
    n = n_;
    L = L_;
    stretch = stretch_;
    gamma = gamma_;
    cbuf.resize(n + 2 * G);
    fbuf.resize(n + 2 * G + 1);
    for (int i = 1 - G; i <= n + G; ++i) cbuf[i - (1 - G)] = coord_centre(i, n, L, stretch, gamma);
    for (int i = -G; i <= n + G; ++i) fbuf[i + G] = coord_face(i, n, L, stretch, gamma);
}

std::string grid_key(const tbfg_grid_desc* g, int dev)
{
    char buf[256];
    std::snprintf(buf, sizeof buf, "%d,%d,%d|%a,%a,%a|%d,%d,%d|%a|dev%d", g->n[0], g->n[1], g->n[2], g->L[0], g->L[1],
                  g->L[2], g->stretch[0], g->stretch[1], g->stretch[2], g->gamma, dev);
    return std::string(buf);
}

int validate_grid(const tbfg_grid_desc* g)
{
    if (!g) {
        set_error("grid descriptor is NULL");
        return TBFG_EINVAL;
    }
    for (int a = 0; a < 3; ++a) {
        if (g->n[a] < 1 || g->n[a] > (1 << 14)) {
            set_error("grid axis %d has %d cells (expected 1..16384)", a, g->n[a]);
            return TBFG_EINVAL;
        }
        if (!(g->L[a] > 0.0) || !std::isfinite(g->L[a])) {
            set_error("grid axis %d has length %g", a, g->L[a]);
            return TBFG_EINVAL;
        }
        if (g->stretch[a] != 0 && g->stretch[a] != 1) {
            set_error("grid axis %d has stretch flag %d (0 uniform, 1 tanh)", a, g->stretch[a]);
            return TBFG_EINVAL;
        }
        if ((a == 0 || a == 2) && g->stretch[a]) {
            set_error("periodic axis %d must be uniform", a);
            return TBFG_EINVAL;
        }
    }
    if (g->stretch[1] && !(g->gamma > 0.0)) {
        set_error("tanh stretching factor must be positive (got %g)", g->gamma);
        return TBFG_EINVAL;
    }
    return TBFG_OK;
}

HostLevel make_level(const tbfg_grid_desc* g, int lev)
{
    HostLevel L;
    for (int a = 0; a < 3; ++a) {
        L.n[a] = g->n[a] >> lev;
        L.ax[a].build(L.n[a], g->L[a], g->stretch[a], g->gamma);
    }
    return L;
}

int DevBuf::alloc(size_t b)
{
    release();
    if (b == 0) b = 16;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
        p = nullptr;
        set_error("hipMalloc(%zu) failed: %s", b, hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? TBFG_ENOMEM : TBFG_EDEVICE;
    }
    bytes = b;
    return TBFG_OK;
}

void DevBuf::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

// one non-blocking stream per device (default priority), kept for the process
static int copy_stream(hipStream_t* out)
{
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static hipStream_t streams[kMaxDevices] = {};
    int dev = 0;
    TBFG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device %d beyond the copy streams' table", dev);
        return TBFG_EDEVICE;
    }
    std::lock_guard<std::mutex> lk(mu);
    if (!streams[dev]) TBFG_HIP(hipStreamCreateWithFlags(&streams[dev], hipStreamNonBlocking));
    *out = streams[dev];
    return TBFG_OK;
}

int copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind)
{
    hipStream_t s;
    TBFG_TRY(copy_stream(&s));
    TBFG_HIP(hipMemcpyAsync(dst, src, bytes, kind, s));
    TBFG_HIP(hipStreamSynchronize(s));
    return TBFG_OK;
}

int memset_sync(void* dst, int value, size_t bytes)
{
    hipStream_t s;
    TBFG_TRY(copy_stream(&s));
    TBFG_HIP(hipMemsetAsync(dst, value, bytes, s));
    TBFG_HIP(hipStreamSynchronize(s));
    return TBFG_OK;
}

int upload(DevBuf& d, const void* src, size_t bytes)
{
    TBFG_TRY(d.alloc(bytes));
    return copy_sync(d.p, src, bytes, hipMemcpyHostToDevice);
}

template <class T>
int LevelTables<T>::build(const HostLevel& L, const HostLevel* C)
{
    // layout per axis: ip[n], im[n], rlo[nc], rhi[nc], plo[n], phi[n]
    size_t total = 0;
    for (int a = 0; a < 3; ++a) total += 4 * (size_t)L.n[a] + 2 * (size_t)(C ? C->n[a] : 0);
    std::vector<T> h(total + 1);
    size_t off[3][6];
    size_t o = 0;
    for (int a = 0; a < 3; ++a) {
        const Axis& f = L.ax[a];
        const int n = L.n[a];
        const int nc = C ? C->n[a] : 0;
        off[a][0] = o;
        for (int i = 1; i <= n; ++i) h[o++] = (T)(1.0 / (f.dc(i + 1) * f.df(i)));
        off[a][1] = o;
        for (int i = 1; i <= n; ++i) h[o++] = (T)(1.0 / (f.dc(i) * f.df(i)));
        off[a][2] = o;
        off[a][3] = o + nc;
        if (C) {
            const Axis& c = C->ax[a];
            for (int I = 1; I <= nc; ++I) {
                const double span = f.c(2 * I) - f.c(2 * I - 1);
                h[o + I - 1] = (T)((f.c(2 * I) - c.c(I)) / span);
                h[o + nc + I - 1] = (T)((c.c(I) - f.c(2 * I - 1)) / span);
            }
        }
        o += 2 * (size_t)nc;
        off[a][4] = o;
        off[a][5] = o + n;
        if (C) {
            const Axis& c = C->ax[a];
            for (int i = 1; i <= n; ++i) {
                const int ih = i / 2;
                const double span = c.c(ih + 1) - c.c(ih);
                h[o + i - 1] = (T)((c.c(ih + 1) - f.c(i)) / span);
                h[o + n + i - 1] = (T)((f.c(i) - c.c(ih)) / span);
            }
        }
        o += 2 * (size_t)n;
    }
    TBFG_TRY(upload(buf, h.data(), h.size() * sizeof(T)));
    const T* base = buf.as<T>();
    view.nx = L.n[0];
    view.ny = L.n[1];
    view.nz = L.n[2];
    for (int a = 0; a < 3; ++a) {
        view.a[a].ip = base + off[a][0];
        view.a[a].im = base + off[a][1];
        view.a[a].rlo = base + off[a][2];
        view.a[a].rhi = base + off[a][3];
        view.a[a].plo = base + off[a][4];
        view.a[a].phi = base + off[a][5];
    }
    return TBFG_OK;
}

template struct LevelTables<double>;
template struct LevelTables<float>;

}  // namespace tbfg
