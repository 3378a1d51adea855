// Runtime behind include/msod2d.h: device selection and stream, the data
// region (host-address-keyed device mirrors, the OpenACC `enter data` model of
// the reference driver, src/app/sources/main.f90:197-425), the element-major
// layouts derived once per mesh, table validation and the kernel calls.
#include "msod2d.h"
#include "msod_closure.h"
#include "msod_gen.h"
#include "msod_geom.h"
#include <rocprofiler-sdk-roctx/roctx.h>

#include "msod_kernels.h"
#include "msod_mlp.h"
#include "msod_plan.h"
#include "msod_tet.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <atomic>
#include <vector>

// Partial sums of msod_closure_gradient_noalloc, which allocates nothing:
// one row of kMlpTrain floats per workgroup of the gradient kernel, at most
// kNoAllocBlocks workgroups (1024 x 1505 floats = 6.2 MB per GPU, held until
// the process ends).  Host code never uses this name as an address: the
// device address of each GPU's copy is taken once per device in msod_init
// (hipGetSymbolAddress) and used from then on.
constexpr int kNoAllocBlocks = 1024;
static __device__ __align__(256) float scratch[kNoAllocBlocks * msod::kMlpTrain];

// roctx range around each entry point: the counterpart of the reference's
// nvtxStartRange/nvtxEndRange inside both kernels
// (src/utils/sources/mod_nvtx.f90:54-80, elem_convec.f90:221,429,
// elem_diffu.f90:34,196), without the device synchronisation those force
// (mod_nvtx.f90:60,77), so the ranges do not change the timing.  Visible in
// rocprofv3 --marker-trace.
struct MsodRange {
    explicit MsodRange(const char *name) { roctxRangePushA(name); }
    ~MsodRange() { roctxRangePop(); }
    MsodRange(const MsodRange &) = delete;
    MsodRange &operator=(const MsodRange &) = delete;
};

namespace msod {

// ---------------------------------------------------------------------------
// layout kernels
// ---------------------------------------------------------------------------
// Zero n doubles with 16-byte stores, one per thread (p is 16-byte aligned:
// hipMalloc bases and whole arrays are passed).  A grid covering the array
// in one trip measured faster than grid-stride loops of 1024 or 4096
// workgroups.
__global__ void k_zero(double *p, int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t n2 = n / 2;
    double2 *q = reinterpret_cast<double2 *>(p);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride)
        q[i] = make_double2(0.0, 0.0);
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) p[n - 1] = 0.0;
}

// The three outputs in one launch (each 16-byte aligned): thread i stores one
// 16-byte zero into the array whose range holds i.  Used below 16 M doubles,
// where launches dominate (8^3 p=3: 0.109 -> 0.101 ms per pair); on the
// 128^3 p=4 outputs three launches of k_zero measured 0.78 against 0.83 ms.
struct ZeroSet {
    double *p[3];
    int64_t n[3];
};

__global__ void k_zero3(ZeroSet z)
{
    const int64_t a = z.n[0] / 2, b = z.n[1] / 2, c = z.n[2] / 2;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double2 *q = nullptr;
    if (i < a)
        q = reinterpret_cast<double2 *>(z.p[0]) + i;
    else if (i < a + b)
        q = reinterpret_cast<double2 *>(z.p[1]) + (i - a);
    else if (i < a + b + c)
        q = reinterpret_cast<double2 *>(z.p[2]) + (i - a - b);
    if (q) *q = make_double2(0.0, 0.0);
    if (i == 0)
        for (int k = 0; k < 3; ++k)
            if (z.n[k] & 1) z.p[k][z.n[k] - 1] = 0.0;
}

__global__ void k_zero_unaligned(double *p, int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0.0;
}

// Zero n floats: 16-byte stores over the 16-byte aligned body, single
// stores for the head and tail.
__global__ void k_zero_f32(float *p, int64_t n)
{
    const int64_t head = std::min<int64_t>(n, (int64_t)((16u - ((uintptr_t)p & 15u)) & 15u) / 4);
    const int64_t body4 = (n - head) / 4;
    const int64_t tail0 = head + body4 * 4;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    float4 *q = reinterpret_cast<float4 *>(p + head);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < body4; i += stride)
        q[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (blockIdx.x == 0 && threadIdx.x < 8) {
        if (threadIdx.x < head) p[threadIdx.x] = 0.f;
        if (tail0 + threadIdx.x < n) p[tail0 + threadIdx.x] = 0.f;
    }
}

template <typename TO>
__global__ void k_zero_nodes(TO *Rmass, TO *Rmom, TO *Rener, int64_t npoin,
                             const uint32_t *nodes, int64_t count)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        const int64_t g = nodes[i];
        Rmass[g] = 0.0;
        Rener[g] = 0.0;
        Rmom[g] = 0.0;
        Rmom[g + npoin] = 0.0;
        Rmom[g + 2 * npoin] = 0.0;
    }
}

// geo[e][k][l][0..1]: the pairs (He c = 2k, 2k+1) for k < 4, (He c = 8,
// gpvol) for k = 4 (GeoNode), from He(3,3,nn,nelem) and gpvol(1,nn,nelem).
template <typename T>
__global__ void k_build_geo(const T *He, const T *gpvol, int64_t nelem, int nn, const int32_t *l2g,
                            T *geo)
{
    const int64_t total = nelem * nn;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const int64_t e = i / nn;
        const int l = (int)(i - e * nn);
        const int64_t a = l2g[l];
        const T *src = He + 9 * (a + (int64_t)nn * e);
        T v[10];
#pragma unroll
        for (int c = 0; c < 9; ++c) v[c] = src[c];
        v[9] = gpvol[a + (int64_t)nn * e];
        T *dst = geo + 2 * (e * 5 * nn + l);
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            dst[2 * k * nn] = v[2 * k];
            dst[2 * k * nn + 1] = v[2 * k + 1];
        }
    }
}

static unsigned grid_for(int64_t n, int bs)
{
    int64_t g = (n + bs - 1) / bs;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, 65536));
}

hipError_t launch_zero_outputs(double *Rmass, double *Rmom, double *Rener, int64_t n, hipStream_t s)
{
    const int64_t len[3] = {n, 3 * n, n};
    double *ptr[3] = {Rmass, Rmom, Rener};
    if ((((uintptr_t)Rmass | (uintptr_t)Rmom | (uintptr_t)Rener) & 15u) == 0 && 5 * n < (int64_t(1) << 24)) {
        const ZeroSet z{{Rmass, Rmom, Rener}, {len[0], len[1], len[2]}};
        const int64_t blocks = std::min<int64_t>((len[0] / 2 + len[1] / 2 + len[2] / 2) / 256 + 1, 0x7fffffff);
        hipLaunchKernelGGL(k_zero3, dim3((unsigned)blocks), dim3(256), 0, s, z);
        return hipGetLastError();
    }
    for (int k = 0; k < 3; ++k) {
        const bool aligned = ((uintptr_t)ptr[k] & 15u) == 0;
        // 16-byte stores need a 16-byte aligned base; otherwise one double at a time
        if (aligned) {
            const int64_t blocks = std::min<int64_t>(len[k] / 2 / 256 + 1, 0x7fffffff);
            hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, s, ptr[k], len[k]);
        } else {
            hipLaunchKernelGGL(k_zero_unaligned, dim3(grid_for(len[k], 256)), dim3(256), 0, s, ptr[k], len[k]);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Zero n doubles at a 16-byte aligned p (a kernel, not a memset: the calls
// stay safe to capture in a HIP graph).
hipError_t launch_zero_doubles(double *p, int64_t n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>(n / 2 / 256 + 1, 0x7fffffff);
    hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

template <typename TO>
hipError_t launch_zero_nodes(TO *Rmass, TO *Rmom, TO *Rener, int64_t npoin, const uint32_t *nodes,
                             int64_t count, hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_nodes<TO>, dim3(grid_for(count, 256)), dim3(256), 0, s, Rmass, Rmom, Rener,
                       npoin, nodes, count);
    return hipGetLastError();
}
template hipError_t launch_zero_nodes<double>(double *, double *, double *, int64_t, const uint32_t *,
                                              int64_t, hipStream_t);
template hipError_t launch_zero_nodes<float>(float *, float *, float *, int64_t, const uint32_t *, int64_t,
                                             hipStream_t);

hipError_t launch_zero_outputs(float *Rmass, float *Rmom, float *Rener, int64_t n, hipStream_t s)
{
    const int64_t len[3] = {n, 3 * n, n};
    float *ptr[3] = {Rmass, Rmom, Rener};
    for (int k = 0; k < 3; ++k) {
        const int64_t blocks = std::min<int64_t>(len[k] / 4 / 256 + 1, 0x7fffffff);
        hipLaunchKernelGGL(k_zero_f32, dim3((unsigned)blocks), dim3(256), 0, s, ptr[k], len[k]);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

int batch_size(int porder)
{
    switch (porder) {
    case 1: return Shape<1>::BATCH;
    case 2: return Shape<2>::BATCH;
    case 3: return Shape<3>::BATCH;
    case 4: return Shape<4>::BATCH;
    case 5: return Shape<5>::BATCH;
    case 6: return Shape<6>::BATCH;
    case 7: return Shape<7>::BATCH;
    case 8: return Shape<8>::BATCH;
    default: return 1;
    }
}

int convec_batch_size(int porder)
{
    switch (porder) {
    case 1: return CShape<1>::BATCH;
    case 2: return CShape<2>::BATCH;
    case 3: return CShape<3>::BATCH;
    case 4: return CShape<4>::BATCH;
    case 5: return CShape<5>::BATCH;
    case 6: return CShape<6>::BATCH;
    case 7: return CShape<7>::BATCH;
    case 8: return CShape<8>::BATCH;
    default: return 1;
    }
}

template <typename T>
hipError_t launch_build_geo(const T *He, const T *gpvol, int64_t nelem, int nn,
                            const int32_t *l2g, T *geo, hipStream_t s)
{
    hipLaunchKernelGGL((k_build_geo<T>), dim3(grid_for(nelem * nn, 256)), dim3(256), 0, s, He,
                       gpvol, nelem, nn, l2g, geo);
    return hipGetLastError();
}
template hipError_t launch_build_geo<double>(const double *, const double *, int64_t, int,
                                             const int32_t *, double *, hipStream_t);
template hipError_t launch_build_geo<float>(const float *, const float *, int64_t, int,
                                            const int32_t *, float *, hipStream_t);

}  // namespace msod

// ---------------------------------------------------------------------------
// runtime state
// ---------------------------------------------------------------------------
namespace {

using namespace msod;

// An array of the data region.  Keyed by its host address (msod_data_create,
// msod_data_copyin) or, for a device array the caller owns and registered
// with msod_data_register, by its device address (owned = false: never
// freed by the library).
struct Region {
    size_t bytes;
    void *dev;
    uint64_t version;
    bool owned = true;
};

struct TableSet {
    int porder;
    std::vector<double> D;       // q*q
    std::vector<int32_t> l2g;    // nn
    double *dD = nullptr;
    int32_t *dl2g = nullptr;
};

// Assembly plan of one mesh (see msod_kernels.h).
struct PlanEntry {
    const void *src = nullptr;
    uint64_t ver = 0;
    int64_t nelem = 0, npoin = 0;
    int porder = 0;
    std::vector<int32_t> l2g;
    int B = 1;
    int64_t nbatch = 0;
    uint32_t *conn = nullptr, *uoff = nullptr, *unode = nullptr, *order = nullptr;
    uint16_t *lidx = nullptr;
    std::vector<int64_t> color_off;  // batches of colour c: order[color_off[c] .. color_off[c+1])
    uint32_t *unref = nullptr;       // nodes no element references (zeroed per call)
    int64_t nunref = 0;
    bool colored = false;            // batch colouring built (the coloured assembly needs it)
    void release()
    {
        for (void *p : {(void *)conn, (void *)uoff, (void *)unode, (void *)order, (void *)lidx,
                        (void *)unref})
            if (p) (void)hipFree(p);
        conn = uoff = unode = order = unref = nullptr;
        lidx = nullptr;
    }
};

struct DerivedGeo {
    const void *srcHe, *srcVol;
    uint64_t verHe, verVol;
    int64_t nelem;
    int porder;
    size_t elt;
    std::vector<int32_t> l2g;
    void *geo;  // He and gpvol, element-major pairs (k_build_geo)
};

// A tetrahedral connectivity found valid (every entry in 1..npoin), with
// how its element residuals reach the nodes: plain stores when no node is
// shared (then the nodes no element references are zeroed from a list),
// atomics into zero-filled outputs otherwise.
struct CheckedMesh {
    const void *src;
    uint64_t ver;
    int64_t n, npoin;
    bool disjoint;
    uint32_t *unref;  // device list of unreferenced nodes (disjoint meshes)
    int64_t nunref;
    void release()
    {
        if (unref) (void)hipFree(unref);
        unref = nullptr;
    }
};

constexpr int kNumTimers = 20;
struct Pending {
    int kernel;
    hipEvent_t a, b;
};

struct State {
    bool init = false;
    int device = 0;
    hipStream_t stream = nullptr;
    bool async = false;
    bool timing = false;
    double c_rho = 1.0, c_ener = 1.0;
    std::map<uintptr_t, Region> present;
    uint64_t next_version = 1;
    std::vector<TableSet *> tables;
    std::vector<PlanEntry> plans;
    std::vector<DerivedGeo> geos;
    std::vector<CheckedMesh> checked;
    int assembly = kAtomic;
    double ms[kNumTimers] = {0};
    int64_t launches[kNumTimers] = {0};
    std::vector<Pending> pending;
    std::vector<hipEvent_t> free_events;
    // library-owned device buffers of the closure calls (grow-only)
    void *scr[4] = {nullptr, nullptr, nullptr, nullptr};
    size_t scr_bytes[4] = {0, 0, 0, 0};
};

State g;
std::recursive_mutex g_mu;
// device address of `scratch` on each device, set by msod_init
constexpr int kMaxDevices = 64;
float *g_noalloc_scratch[kMaxDevices] = {};
thread_local std::string g_err;

int fail(int status, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return status;
}

int hipfail(hipError_t e, const char *what)
{
    return fail(MSOD_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

#define HIPCHK(call, what)                          \
    do {                                            \
        hipError_t e_ = (call);                     \
        if (e_ != hipSuccess) return hipfail(e_, what); \
    } while (0)

#define NEED_INIT()                                                         \
    do {                                                                    \
        if (!g.init) return fail(MSOD_ERR_STATE, "msod_init has not been called"); \
    } while (0)

// Present-table lookup by containment.
Region *find_region(const void *p, uintptr_t *base)
{
    if (g.present.empty()) return nullptr;
    auto it = g.present.upper_bound((uintptr_t)p);
    if (it == g.present.begin()) return nullptr;
    --it;
    if ((uintptr_t)p < it->first + it->second.bytes) {
        *base = it->first;
        return &it->second;
    }
    return nullptr;
}

// Region whose DEVICE copy holds address p (a present array reached through
// msod_data_deviceptr, or a registered device array).
Region *find_region_dev(const void *p)
{
    for (auto &kv : g.present) {
        const uintptr_t lo = (uintptr_t)kv.second.dev;
        if ((uintptr_t)p >= lo && (uintptr_t)p < lo + kv.second.bytes) return &kv.second;
    }
    return nullptr;
}

bool is_device_pointer(const void *p)
{
    hipPointerAttribute_t attr;
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice;
}

// Host or device address of an array argument -> device address.
int resolve(const void *p, size_t need, const char *name, const void **dev, uint64_t *ver)
{
    if (!p) return fail(MSOD_ERR_ARG, "%s is NULL", name);
    uintptr_t base = 0;
    if (Region *r = find_region(p, &base)) {
        const size_t off = (uintptr_t)p - base;
        if (off + need > r->bytes)
            return fail(MSOD_ERR_MESH, "%s: present array holds %zu bytes past this address, %zu needed",
                        name, r->bytes - off, need);
        *dev = (const char *)r->dev + off;
        *ver = r->version;
        return MSOD_OK;
    }
    if (Region *r = find_region_dev(p)) {  // the device address of a present array
        const size_t off = (uintptr_t)p - (uintptr_t)r->dev;
        if (off + need > r->bytes)
            return fail(MSOD_ERR_MESH, "%s: present array holds %zu bytes past this address, %zu needed",
                        name, r->bytes - off, need);
        *dev = p;
        *ver = r->version;
        return MSOD_OK;
    }
    if (is_device_pointer(p)) {
        // device memory outside the data region: version 0, and every layout
        // derived from it is rebuilt on each call (the caller may change it
        // without telling the library; msod_data_register caches them)
        *dev = p;
        *ver = 0;
        return MSOD_OK;
    }
    return fail(MSOD_ERR_NOT_PRESENT, "%s (%p) is not present on the device", name, p);
}

// Small tables are read on the host: directly when the address is host
// memory (present or not), through a copy when it is device memory.
int host_view(const void *p, size_t bytes, std::vector<char> &tmp, const char **out, const char *name)
{
    if (!p) return fail(MSOD_ERR_ARG, "%s is NULL", name);
    if (is_device_pointer(p)) {
        tmp.resize(bytes);
        HIPCHK(hipMemcpy(tmp.data(), p, bytes, hipMemcpyDeviceToHost), "reading a device table");
        *out = tmp.data();
    } else {
        *out = (const char *)p;
    }
    return MSOD_OK;
}

// Checks invAtoIJK/AtoI/J/K/dlxigp_ip and returns the cached device table set.
// dl32: dlxigp_ip holds float (a caller built at rp = 4); its values are
// widened to double exactly.
int get_tables(int porder, const int32_t *invAtoIJK, const int32_t *AtoI, const int32_t *AtoJ,
               const int32_t *AtoK, const void *dlxigp_ip, TableSet **out, bool dl32 = false)
{
    const int q = porder + 1, nn = q * q * q;
    std::vector<char> t0, t1, t2, t3, t4;
    const char *pinv, *pi, *pj, *pk, *pdl;
    int st;
    if ((st = host_view(invAtoIJK, sizeof(int32_t) * nn, t0, &pinv, "invAtoIJK"))) return st;
    if ((st = host_view(AtoI, sizeof(int32_t) * nn, t1, &pi, "gmshAtoI"))) return st;
    if ((st = host_view(AtoJ, sizeof(int32_t) * nn, t2, &pj, "gmshAtoJ"))) return st;
    if ((st = host_view(AtoK, sizeof(int32_t) * nn, t3, &pk, "gmshAtoK"))) return st;
    if ((st = host_view(dlxigp_ip, (dl32 ? sizeof(float) : sizeof(double)) * nn * 3 * q, t4, &pdl,
                        "dlxigp_ip")))
        return st;
    const int32_t *inv = (const int32_t *)pinv, *ai = (const int32_t *)pi,
                  *aj = (const int32_t *)pj, *ak = (const int32_t *)pk;
    std::vector<double> dlw;
    if (dl32) {
        dlw.resize((size_t)nn * 3 * q);
        for (size_t i = 0; i < dlw.size(); ++i) dlw[i] = (double)((const float *)pdl)[i];
    }
    const double *dl = dl32 ? dlw.data() : (const double *)pdl;

    std::vector<int32_t> l2g(nn);
    std::vector<char> seen(nn, 0);
    for (int l = 0; l < nn; ++l) {
        const int32_t a = inv[l];
        if (a < 1 || a > nn || seen[a - 1])
            return fail(MSOD_ERR_TABLE, "invAtoIJK is not a permutation of 1..%d", nn);
        seen[a - 1] = 1;
        l2g[l] = a - 1;
    }
    for (int a = 0; a < nn; ++a) {
        const int i = ai[a], j = aj[a], k = ak[a];
        if (i < 1 || i > q || j < 1 || j > q || k < 1 || k > q ||
            inv[(i - 1) + q * ((j - 1) + q * (k - 1))] != a + 1)
            return fail(MSOD_ERR_TABLE, "gmshAtoI/J/K disagree with invAtoIJK at node %d", a + 1);
    }
    // D[i][n] from the nodes on the xi-line (i,1,1); every node/direction
    // must carry the same row of the 1-D derivative (collocation at GLL points)
    std::vector<double> D(q * q);
    for (int i = 0; i < q; ++i) {
        const int a = inv[i] - 1;
        for (int n = 0; n < q; ++n) D[i * q + n] = dl[a + nn * (0 + 3 * n)];
    }
    for (int a = 0; a < nn; ++a) {
        const int idx[3] = {ai[a] - 1, aj[a] - 1, ak[a] - 1};
        for (int d = 0; d < 3; ++d)
            for (int n = 0; n < q; ++n)
                if (std::memcmp(&dl[a + nn * (d + 3 * n)], &D[idx[d] * q + n], sizeof(double)) != 0)
                    return fail(MSOD_ERR_TABLE,
                                "dlxigp_ip(%d,%d,%d) is not the 1-D derivative row of its node", a + 1,
                                d + 1, n + 1);
    }
    for (TableSet *ts : g.tables)
        if (ts->porder == porder && ts->l2g == l2g && ts->D.size() == D.size() &&
            std::memcmp(ts->D.data(), D.data(), sizeof(double) * q * q) == 0) {
            *out = ts;
            return MSOD_OK;
        }
    TableSet *ts = new TableSet;
    ts->porder = porder;
    ts->D = D;
    ts->l2g = l2g;
    HIPCHK(hipMalloc(&ts->dD, sizeof(double) * q * q), "hipMalloc tables");
    HIPCHK(hipMalloc(&ts->dl2g, sizeof(int32_t) * nn), "hipMalloc tables");
    HIPCHK(hipMemcpy(ts->dD, D.data(), sizeof(double) * q * q, hipMemcpyHostToDevice), "upload D");
    HIPCHK(hipMemcpy(ts->dl2g, l2g.data(), sizeof(int32_t) * nn, hipMemcpyHostToDevice), "upload l2g");
    g.tables.push_back(ts);
    *out = ts;
    return MSOD_OK;
}

constexpr size_t kMaxDerived = 4;

template <class V>
int upload(const std::vector<V> &h, V **d, const char *what)
{
    if (h.empty()) {
        *d = nullptr;
        return MSOD_OK;
    }
    HIPCHK(hipMalloc(d, sizeof(V) * h.size()), what);
    HIPCHK(hipMemcpy(*d, h.data(), sizeof(V) * h.size(), hipMemcpyHostToDevice), what);
    return MSOD_OK;
}

// Builds the assembly plan of a mesh (msod_plan.h) from its connectivity,
// a device array in the reference layout, and uploads it.
int build_plan(const void *dconnec, int64_t nelem, int64_t npoin, const TableSet *ts, int B, PlanEntry &P)
{
    const int64_t E = nelem * (int64_t)ts->l2g.size();
    std::vector<int32_t> raw(E);
    HIPCHK(hipMemcpy(raw.data(), dconnec, sizeof(int32_t) * E, hipMemcpyDeviceToHost),
           "reading connec");
    HostPlan H;
    std::string err;
    if (!build_host_plan(raw.data(), nelem, npoin, ts->l2g, B, H, &err))
        return fail(MSOD_ERR_MESH, "%s", err.c_str());
    std::vector<int32_t>().swap(raw);
    P.B = H.B;
    P.nbatch = H.nbatch;
    P.color_off = H.color_off;
    P.colored = H.colored;
    int st;
    if ((st = upload(H.conn, &P.conn, "plan upload"))) return st;
    if ((st = upload(H.lidx, &P.lidx, "plan upload"))) return st;
    if ((st = upload(H.uoff, &P.uoff, "plan upload"))) return st;
    if ((st = upload(H.unode, &P.unode, "plan upload"))) return st;
    if ((st = upload(H.order, &P.order, "plan upload"))) return st;
    if ((st = upload(H.unref, &P.unref, "plan upload"))) return st;
    P.nunref = (int64_t)H.unref.size();
    return MSOD_OK;
}

// Layouts derived from device memory outside the data region (version 0) are
// not reused from call to call: the array may have changed since.
void drop_unversioned(const void *src)
{
    for (size_t i = g.plans.size(); i-- > 0;)
        if (g.plans[i].src == src && g.plans[i].ver == 0) {
            g.plans[i].release();
            g.plans.erase(g.plans.begin() + i);
        }
    for (size_t i = g.checked.size(); i-- > 0;)
        if (g.checked[i].src == src && g.checked[i].ver == 0) {
            g.checked[i].release();
            g.checked.erase(g.checked.begin() + i);
        }
    for (size_t i = g.geos.size(); i-- > 0;)
        if ((g.geos[i].srcHe == src && g.geos[i].verHe == 0) ||
            (g.geos[i].srcVol == src && g.geos[i].verVol == 0)) {
            (void)hipFree(g.geos[i].geo);
            g.geos.erase(g.geos.begin() + i);
        }
}

int get_plan(const void *dconnec, uint64_t ver, int64_t nelem, int64_t npoin, const TableSet *ts, int B,
             const PlanEntry **out)
{
    if (ver == 0) drop_unversioned(dconnec);
    for (auto &c : g.plans)
        if (c.src == dconnec && c.ver == ver && c.nelem == nelem && c.npoin == npoin &&
            c.porder == ts->porder && c.l2g == ts->l2g && c.B == B) {
            *out = &c;
            return MSOD_OK;
        }
    PlanEntry P;
    P.src = dconnec;
    P.ver = ver;
    P.nelem = nelem;
    P.npoin = npoin;
    P.porder = ts->porder;
    P.l2g = ts->l2g;
    int st = build_plan(dconnec, nelem, npoin, ts, B, P);
    if (st) {
        P.release();
        return st;
    }
    if (g.plans.size() >= kMaxDerived) {
        g.plans.front().release();
        g.plans.erase(g.plans.begin());
    }
    g.plans.push_back(std::move(P));
    *out = &g.plans.back();
    return MSOD_OK;
}

template <typename T>
int get_geo(const void *dHe, uint64_t vHe, const void *dVol, uint64_t vVol, int64_t nelem,
            const TableSet *ts, const T **geo)
{
    if (vHe == 0) drop_unversioned(dHe);
    if (vVol == 0) drop_unversioned(dVol);
    for (auto &c : g.geos)
        if (c.srcHe == dHe && c.srcVol == dVol && c.verHe == vHe && c.verVol == vVol &&
            c.nelem == nelem && c.porder == ts->porder && c.elt == sizeof(T) && c.l2g == ts->l2g) {
            *geo = (const T *)c.geo;
            return MSOD_OK;
        }
    const int nn = (int)ts->l2g.size();
    DerivedGeo c{dHe, dVol, vHe, vVol, nelem, ts->porder, sizeof(T), ts->l2g, nullptr};
    HIPCHK(hipMalloc(&c.geo, sizeof(T) * 10 * nelem * nn), "hipMalloc element-major He and gpvol");
    hipError_t e = launch_build_geo<T>((const T *)dHe, (const T *)dVol, nelem, nn, ts->dl2g, (T *)c.geo,
                                       g.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g.stream);
    if (e != hipSuccess) {
        (void)hipFree(c.geo);
        return hipfail(e, "building element-major geometry");
    }
    if (g.geos.size() >= kMaxDerived) {
        (void)hipFree(g.geos.front().geo);
        g.geos.erase(g.geos.begin());
    }
    g.geos.push_back(c);
    *geo = (const T *)c.geo;
    return MSOD_OK;
}

void drop_derived_in(uintptr_t lo, uintptr_t hi)
{
    auto inside = [&](const void *p) { return (uintptr_t)p >= lo && (uintptr_t)p < hi; };
    for (size_t i = g.plans.size(); i-- > 0;)
        if (inside(g.plans[i].src)) {
            g.plans[i].release();
            g.plans.erase(g.plans.begin() + i);
        }
    for (size_t i = g.checked.size(); i-- > 0;)
        if (inside(g.checked[i].src)) {
            g.checked[i].release();
            g.checked.erase(g.checked.begin() + i);
        }
    for (size_t i = g.geos.size(); i-- > 0;)
        if (inside(g.geos[i].srcHe) || inside(g.geos[i].srcVol)) {
            (void)hipFree(g.geos[i].geo);
            g.geos.erase(g.geos.begin() + i);
        }
}

// The contents of an array changed (the stream is synchronised): a present
// array takes a new version, and every layout derived from its device copy
// is dropped, also those cached for calls that passed the device address
// itself (version 0), which a new version alone would leave in use.
void contents_changed(const void *ptr)
{
    uintptr_t base;
    Region *r = find_region(ptr, &base);
    if (!r) r = find_region_dev(ptr);
    if (r) {
        r->version = g.next_version++;
        drop_derived_in((uintptr_t)r->dev, (uintptr_t)r->dev + r->bytes);
    } else {
        drop_derived_in((uintptr_t)ptr, (uintptr_t)ptr + 1);
    }
}

// A device routine wrote the array at ptr (host key or device address).
void bump_version(const void *ptr)
{
    uintptr_t base;
    Region *r = find_region(ptr, &base);
    if (!r) r = find_region_dev(ptr);
    if (r) r->version = g.next_version++;
}

// --- timing ---------------------------------------------------------------
hipEvent_t take_event()
{
    if (!g.free_events.empty()) {
        hipEvent_t e = g.free_events.back();
        g.free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

struct TimedLaunch {
    int kernel;
    hipEvent_t a = nullptr;
    explicit TimedLaunch(int k) : kernel(k)
    {
        if (g.timing) {
            a = take_event();
            (void)hipEventRecord(a, g.stream);
        }
    }
    void done()
    {
        if (!a) return;
        hipEvent_t b = take_event();
        (void)hipEventRecord(b, g.stream);
        g.pending.push_back({kernel, a, b});
        a = nullptr;
    }
};

void collect_timing()
{
    for (auto &p : g.pending) {
        float ms = 0.f;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            g.ms[p.kernel] += ms;
            g.launches[p.kernel] += 1;
        }
        g.free_events.push_back(p.a);
        g.free_events.push_back(p.b);
    }
    g.pending.clear();
}

int finish_call(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hipfail(e, what);
    if (!g.async) {
        e = hipStreamSynchronize(g.stream);
        if (e != hipSuccess) return hipfail(e, what);
    }
    if (g.pending.size() > 4096) collect_timing();
    return MSOD_OK;
}

int check_sizes(int porder, int nelem, int npoin)
{
    if (porder < 1 || porder > kMaxP)
        return fail(MSOD_ERR_PORDER, "porder=%d: kernels are compiled for 1..%d", porder, kMaxP);
    if (nelem <= 0 || npoin <= 0) return fail(MSOD_ERR_ARG, "nelem=%d npoin=%d", nelem, npoin);
    return MSOD_OK;
}

// Batch order of the atomic mode, per kernel (measured at 128^3 porder 4):
// mesh order lets neighbouring batches meet in L2, which pays in the
// diffusion; the convection ran faster in colour order (12.81 against 13.30
// ms; colour order within chunks of 256..16384 batches: 13.07..19.90 ms).
constexpr bool kConvecNatural = false;
constexpr bool kDiffuNatural = true;
// Batches per workgroup, at most, per kernel.  The diffusion takes one: a
// form that ran kpb batches per workgroup with every load one pass ahead
// needed 148-152 VGPRs against 120 (3 waves per SIMD against 4) and ran
// slower (128^3 p=4: 14.4-14.9 against 13.4 ms for kpb 2..16; 48^3 p=7: 5.2
// against 3.7 ms).
constexpr int kConvecKpb = kBatchesPerBlock;
constexpr int kDiffuKpb = 1;

// Zero-fill (atomic mode) or unreferenced-node zeroing (coloured mode), then
// the kernel: one launch per colour, or one launch over all batches, taken
// in mesh order (natural) or grouped by colour.
template <class A, class L>
int run_assembled(A &a, const PlanEntry &P, int timer, L launch, const char *what, bool natural, int kpb_max)
{
    Plan pl;
    pl.conn = P.conn;
    pl.lidx = P.lidx;
    pl.uoff = P.uoff;
    pl.unode = P.unode;
    pl.mode = g.assembly;
    // batches per workgroup: kpb_max on large meshes, fewer where that would
    // leave the launch short of 2048 workgroups (8 per CU); 8^3 p=3 has 64
    // batches, which 8 per workgroup put on 8 CUs
    auto kpb_for = [kpb_max](int64_t nb) {
        return (int)std::max<int64_t>(1, std::min<int64_t>(kpb_max, nb / 2048));
    };
    const int64_t NP = a.npoin;
    if (g.assembly == kColored && !P.colored)
        return fail(MSOD_ERR_MESH, "%s: the batches of this mesh need more than 32 colours (elements in no "
                                   "spatial order); the coloured assembly is not available, the atomic one "
                                   "(MSOD_ASSEMBLY_ATOMIC, the default) handles any element order", what);
    {
        TimedLaunch tz(0);
        if (g.assembly == kAtomic)
            HIPCHK(launch_zero_outputs(a.Rmass, a.Rmom, a.Rener, NP, g.stream), "zero-fill launch");
        else
            HIPCHK(launch_zero_nodes(a.Rmass, a.Rmom, a.Rener, NP, P.unref, P.nunref, g.stream),
                   "zero-fill launch");
        tz.done();
    }
    TimedLaunch tk(timer);
    if (g.assembly == kAtomic) {
        pl.batches = natural ? nullptr : P.order;  // null: mesh order, no list
        pl.nb = P.nbatch;
        pl.kpb = kpb_for(pl.nb);
        a.plan = pl;
        HIPCHK(launch(a), what);
    } else {
        for (size_t c = 0; c + 1 < P.color_off.size(); ++c) {
            pl.batches = P.order + P.color_off[c];
            pl.nb = P.color_off[c + 1] - P.color_off[c];
            pl.kpb = kpb_for(pl.nb);
            a.plan = pl;
            HIPCHK(launch(a), what);
        }
    }
    tk.done();
    return MSOD_OK;
}

// Shared front half of the convective calls (TG: geometry, TF: fields, TO:
// outputs; an FP32 derivative table comes with FP32 outputs).
template <typename TG, typename TF, typename TO = double, typename TU = TF, typename TP = TF>
int convec_common(int porder, int nelem, int npoin, const int32_t *connec, const TG *He,
                  const TG *gpvol, const void *dlxigp_ip, const int32_t *invAtoIJK,
                  const int32_t *AtoI, const int32_t *AtoJ, const int32_t *AtoK, const TU *u,
                  const TF *q, const TF *rho, const TP *pr, const TF *E, double gamma_gas,
                  TO *Rmass, TO *Rmom, TO *Rener, int timer, bool fp32_arith = false)
{
    NEED_INIT();
    int st = check_sizes(porder, nelem, npoin);
    if (st) return st;
    const int qn = porder + 1;
    const int64_t nn = (int64_t)qn * qn * qn, NE = nelem, NP = npoin;
    TableSet *ts;
    if ((st = get_tables(porder, invAtoIJK, AtoI, AtoJ, AtoK, dlxigp_ip, &ts, sizeof(TO) == 4))) return st;
    const void *dc, *dHe, *dV, *du, *dq, *dr, *dp = nullptr, *dE, *dRm, *dRmo, *dRe;
    uint64_t vc, vHe, vV, vx;
    if ((st = resolve(connec, sizeof(int32_t) * NE * nn, "connec", &dc, &vc))) return st;
    if ((st = resolve(He, sizeof(TG) * 9 * NE * nn, "He", &dHe, &vHe))) return st;
    if ((st = resolve(gpvol, sizeof(TG) * NE * nn, "gpvol", &dV, &vV))) return st;
    if ((st = resolve(u, sizeof(TU) * 3 * NP, "u", &du, &vx))) return st;
    if ((st = resolve(q, sizeof(TF) * 3 * NP, "q", &dq, &vx))) return st;
    if ((st = resolve(rho, sizeof(TF) * NP, "rho", &dr, &vx))) return st;
    if (pr && (st = resolve(pr, sizeof(TP) * NP, "pr", &dp, &vx))) return st;
    if ((st = resolve(E, sizeof(TF) * NP, "E", &dE, &vx))) return st;
    if ((st = resolve(Rmass, sizeof(TO) * NP, "Rmass", &dRm, &vx))) return st;
    if ((st = resolve(Rmom, sizeof(TO) * 3 * NP, "Rmom", &dRmo, &vx))) return st;
    if ((st = resolve(Rener, sizeof(TO) * NP, "Rener", &dRe, &vx))) return st;
    const PlanEntry *plan;
    const TG *geo;
    if ((st = get_plan(dc, vc, NE, NP, ts, convec_batch_size(porder), &plan))) return st;
    if ((st = get_geo<TG>(dHe, vHe, dV, vV, NE, ts, &geo))) return st;

    ConvecArgs<TG, TF, TO, TU, TP> a;
    a.nelem = NE;
    a.npoin = NP;
    a.geo = geo;
    a.u = (const TU *)du;
    a.q = (const TF *)dq;
    a.rho = (const TF *)dr;
    a.pr = (const TP *)dp;
    a.E = (const TF *)dE;
    a.gamma_m1 = gamma_gas - 1.0;
    a.Rmass = (TO *)dRm;
    a.Rmom = (TO *)dRmo;
    a.Rener = (TO *)dRe;
    a.tab = Tables{ts->dD, ts->dl2g};
    if ((st = run_assembled(a, *plan, timer,
                            [&](const ConvecArgs<TG, TF, TO, TU, TP> &x) {
                                if constexpr (std::is_same<ConvecArgs<TG, TF, TO, TU, TP>,
                                                           ConvecArgs<float, float>>::value)
                                    if (fp32_arith) return launch_convec_fp32(porder, x, g.stream);
                                return launch_convec(porder, x, g.stream);
                            },
                            "convection kernel launch", kConvecNatural, kConvecKpb)))
        return st;
    return finish_call("convection kernel");
}

template <typename TG, typename TF, typename TO = double, typename TU = TF>
int diffu_common(int porder, int nelem, int npoin, const int32_t *connec, const TG *He,
                 const TG *gpvol, const void *dlxigp_ip, const int32_t *invAtoIJK,
                 const int32_t *AtoI, const int32_t *AtoJ, const int32_t *AtoK, double Cp,
                 double Pr, const TF *rho, const TU *u, const TF *Tem, const TF *mu_fluid,
                 const TG *mu_e, const TG *mu_sgs, TO *Rmass, TO *Rmom, TO *Rener,
                 int timer, bool closure = false, const float *dmlp = nullptr, bool fp32_arith = false)
{
    NEED_INIT();
    int st = check_sizes(porder, nelem, npoin);
    if (st) return st;
    const int qn = porder + 1;
    const int64_t nn = (int64_t)qn * qn * qn, NE = nelem, NP = npoin;
    TableSet *ts;
    if ((st = get_tables(porder, invAtoIJK, AtoI, AtoJ, AtoK, dlxigp_ip, &ts, sizeof(TO) == 4))) return st;
    const void *dc, *dHe, *dV, *dr, *du, *dT, *dmf, *dme, *dms = nullptr, *dRm, *dRmo, *dRe;
    uint64_t vc, vHe, vV, vx;
    if ((st = resolve(connec, sizeof(int32_t) * NE * nn, "connec", &dc, &vc))) return st;
    if ((st = resolve(He, sizeof(TG) * 9 * NE * nn, "He", &dHe, &vHe))) return st;
    if ((st = resolve(gpvol, sizeof(TG) * NE * nn, "gpvol", &dV, &vV))) return st;
    if ((st = resolve(rho, sizeof(TF) * NP, "rho", &dr, &vx))) return st;
    if ((st = resolve(u, sizeof(TU) * 3 * NP, "u", &du, &vx))) return st;
    if ((st = resolve(Tem, sizeof(TF) * NP, "Tem", &dT, &vx))) return st;
    if ((st = resolve(mu_fluid, sizeof(TF) * NP, "mu_fluid", &dmf, &vx))) return st;
    if ((st = resolve(mu_e, sizeof(TG) * NE * nn, "mu_e", &dme, &vx))) return st;
    if (!closure && (st = resolve(mu_sgs, sizeof(TG) * NE * nn, "mu_sgs", &dms, &vx))) return st;
    if ((st = resolve(Rmass, sizeof(TO) * NP, "Rmass", &dRm, &vx))) return st;
    if ((st = resolve(Rmom, sizeof(TO) * 3 * NP, "Rmom", &dRmo, &vx))) return st;
    if ((st = resolve(Rener, sizeof(TO) * NP, "Rener", &dRe, &vx))) return st;
    const PlanEntry *plan;
    const TG *geo;
    if ((st = get_plan(dc, vc, NE, NP, ts, batch_size(porder), &plan))) return st;
    if ((st = get_geo<TG>(dHe, vHe, dV, vV, NE, ts, &geo))) return st;

    DiffuArgs<TG, TF, TO, TU> a;
    a.nelem = NE;
    a.npoin = NP;
    a.geo = geo;
    a.rho = (const TF *)dr;
    a.u = (const TU *)du;
    a.Tem = (const TF *)dT;
    a.mu_fluid = (const TF *)dmf;
    a.mu_e = (const TG *)dme;
    a.mu_sgs = (const TG *)dms;
    a.Cp = Cp;
    a.Pr = Pr;
    a.c_rho = g.c_rho;
    a.c_ener = g.c_ener;
    a.Rmass = (TO *)dRm;
    a.Rmom = (TO *)dRmo;
    a.Rener = (TO *)dRe;
    a.tab = Tables{ts->dD, ts->dl2g};
    if ((st = run_assembled(a, *plan, timer,
                            [&](const DiffuArgs<TG, TF, TO, TU> &x) {
                                using AT = DiffuArgs<TG, TF, TO, TU>;
                                if constexpr (std::is_same<AT, DiffuArgs<double, double>>::value)
                                    if (dmlp) return launch_diffu_closure(porder, x, dmlp, g.stream);
                                if constexpr (std::is_same<AT, DiffuArgs<float, float>>::value)
                                    if (fp32_arith) return launch_diffu_fp32(porder, x, g.stream);
                                return launch_diffu(porder, x, g.stream);
                            },
                            "diffusion kernel launch", kDiffuNatural, kDiffuKpb)))
        return st;
    return finish_call("diffusion kernel");
}

// --- learned closure -----------------------------------------------------
enum { kScrWeights = 0, kScrPartial = 1, kScrLoss = 2 };

// Library-owned device buffer `slot` of at least `bytes` (grow-only).
int library_buffer(int slot, size_t bytes, void **p)
{
    if (g.scr_bytes[slot] < bytes) {
        HIPCHK(hipStreamSynchronize(g.stream), "closure buffer resize");
        if (g.scr[slot]) (void)hipFree(g.scr[slot]);
        g.scr[slot] = nullptr;
        g.scr_bytes[slot] = 0;
        HIPCHK(hipMalloc(&g.scr[slot], bytes), "hipMalloc closure buffer");
        g.scr_bytes[slot] = bytes;
    }
    *p = g.scr[slot];
    return MSOD_OK;
}

// Device address of a weight block (present, device, or plain host memory,
// which is copied into a library buffer) and whether the model is usable:
// every value finite and the output scale >= 0.  Reads the values on the host.
int closure_weights(const float *params, const float **dev, bool *usable)
{
    *usable = false;
    *dev = nullptr;
    if (!params) return MSOD_OK;
    const size_t bytes = sizeof(float) * kMlpParams;
    std::vector<float> h(kMlpParams);
    uintptr_t base;
    if (find_region(params, &base) || is_device_pointer(params)) {
        const void *d;
        uint64_t ver;
        int st = resolve(params, bytes, "mlp_params", &d, &ver);
        if (st) return st;
        HIPCHK(hipMemcpyAsync(h.data(), d, bytes, hipMemcpyDeviceToHost, g.stream), "reading mlp_params");
        HIPCHK(hipStreamSynchronize(g.stream), "reading mlp_params");
        *dev = (const float *)d;
    } else {
        std::memcpy(h.data(), params, bytes);
        void *buf;
        int st = library_buffer(kScrWeights, bytes, &buf);
        if (st) return st;
        HIPCHK(hipMemcpyAsync(buf, h.data(), bytes, hipMemcpyHostToDevice, g.stream), "uploading mlp_params");
        HIPCHK(hipStreamSynchronize(g.stream), "uploading mlp_params");
        *dev = (const float *)buf;
    }
    bool ok = h[kOScale] >= 0.f;
    for (float v : h) ok = ok && std::isfinite(v);
    *usable = ok;
    return MSOD_OK;
}

// Device copy of the l -> Gmsh node map of invAtoIJK (cached with the tables).
int get_l2g(int porder, const int32_t *invAtoIJK, const int32_t **dl2g)
{
    const int q = porder + 1, nn = q * q * q;
    std::vector<char> tmp;
    const char *pinv;
    int st = host_view(invAtoIJK, sizeof(int32_t) * nn, tmp, &pinv, "invAtoIJK");
    if (st) return st;
    const int32_t *inv = (const int32_t *)pinv;
    std::vector<int32_t> l2g(nn);
    std::vector<char> seen(nn, 0);
    for (int l = 0; l < nn; ++l) {
        const int32_t a = inv[l];
        if (a < 1 || a > nn || seen[a - 1])
            return fail(MSOD_ERR_TABLE, "invAtoIJK is not a permutation of 1..%d", nn);
        seen[a - 1] = 1;
        l2g[l] = a - 1;
    }
    for (TableSet *ts : g.tables)
        if (ts->porder == porder && ts->l2g == l2g) {
            *dl2g = ts->dl2g;
            return MSOD_OK;
        }
    TableSet *ts = new TableSet;  // map only: D stays empty
    ts->porder = porder;
    ts->l2g = l2g;
    g.tables.push_back(ts);
    HIPCHK(hipMalloc(&ts->dl2g, sizeof(int32_t) * nn), "hipMalloc tables");
    HIPCHK(hipMemcpy(ts->dl2g, l2g.data(), sizeof(int32_t) * nn, hipMemcpyHostToDevice), "upload l2g");
    *dl2g = ts->dl2g;
    return MSOD_OK;
}

// Shared front half of the closure's element kernels: tables, arrays, plan
// (for the 0-based connectivity) and element-major geometry.
int closure_elem_common(int porder, int nelem, int npoin, const int32_t *connec, const double *He,
                        const double *gpvol, const double *dlxigp_ip, const int32_t *invAtoIJK,
                        const int32_t *AtoI, const int32_t *AtoJ, const int32_t *AtoK,
                        const double *rho, const double *u, ClosureElemArgs &a)
{
    NEED_INIT();
    int st = check_sizes(porder, nelem, npoin);
    if (st) return st;
    const int qn = porder + 1;
    const int64_t nn = (int64_t)qn * qn * qn, NE = nelem, NP = npoin;
    TableSet *ts;
    if ((st = get_tables(porder, invAtoIJK, AtoI, AtoJ, AtoK, dlxigp_ip, &ts))) return st;
    const void *dc, *dHe, *dV, *dr, *du;
    uint64_t vc, vHe, vV, vx;
    if ((st = resolve(connec, sizeof(int32_t) * NE * nn, "connec", &dc, &vc))) return st;
    if ((st = resolve(He, sizeof(double) * 9 * NE * nn, "He", &dHe, &vHe))) return st;
    if ((st = resolve(gpvol, sizeof(double) * NE * nn, "gpvol", &dV, &vV))) return st;
    if ((st = resolve(rho, sizeof(double) * NP, "rho", &dr, &vx))) return st;
    if ((st = resolve(u, sizeof(double) * 3 * NP, "u", &du, &vx))) return st;
    const PlanEntry *plan;
    const double *geo;
    if ((st = get_plan(dc, vc, NE, NP, ts, batch_size(porder), &plan))) return st;
    if ((st = get_geo<double>(dHe, vHe, dV, vV, NE, ts, &geo))) return st;
    a = ClosureElemArgs{};
    a.nelem = NE;
    a.npoin = NP;
    a.conn = plan->conn;
    a.geo = geo;
    a.rho = (const double *)dr;
    a.u = (const double *)du;
    a.tab = Tables{ts->dD, ts->dl2g};
    return MSOD_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int msod_init(int device)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    if (g.init) {
        if (device == g.device) return MSOD_OK;
        return fail(MSOD_ERR_STATE, "already initialised on device %d", g.device);
    }
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (device < 0 || device >= n) return fail(MSOD_ERR_ARG, "device %d of %d", device, n);
    HIPCHK(hipSetDevice(device), "hipSetDevice");
    if (device < kMaxDevices && !g_noalloc_scratch[device]) {
        void *ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_noalloc_scratch[device] = (float *)ptr;
    }
    HIPCHK(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking), "hipStreamCreate");
    g.device = device;
    g.init = true;
    return MSOD_OK;
}

int msod_finalize(void)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    if (!g.init) return MSOD_OK;
    (void)hipStreamSynchronize(g.stream);
    collect_timing();
    for (auto &kv : g.present)
        if (kv.second.owned) (void)hipFree(kv.second.dev);
    g.present.clear();
    drop_derived_in(0, UINTPTR_MAX);
    for (TableSet *ts : g.tables) {
        (void)hipFree(ts->dD);
        (void)hipFree(ts->dl2g);
        delete ts;
    }
    g.tables.clear();
    for (hipEvent_t e : g.free_events) (void)hipEventDestroy(e);
    g.free_events.clear();
    for (void *p : g.scr)
        if (p) (void)hipFree(p);
    (void)hipStreamDestroy(g.stream);
    g = State{};
    return MSOD_OK;
}

const char *msod_status_string(int status)
{
    switch (status) {
    case MSOD_OK: return "ok";
    case MSOD_ERR_ARG: return "invalid argument";
    case MSOD_ERR_NOT_PRESENT: return "array not present on the device";
    case MSOD_ERR_PORDER: return "porder not compiled";
    case MSOD_ERR_TABLE: return "inconsistent element tables";
    case MSOD_ERR_HIP: return "HIP runtime error";
    case MSOD_ERR_STATE: return "library not initialised";
    case MSOD_ERR_MESH: return "invalid mesh arrays";
    case MSOD_ERR_MODEL: return "invalid closure model";
    default: return "unknown status";
    }
}

const char *msod_last_error(void) { return g_err.c_str(); }

int msod_data_create(const void *host, size_t bytes)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    NEED_INIT();
    if (!host || bytes == 0) return fail(MSOD_ERR_ARG, "data_create: NULL or empty");
    uintptr_t base;
    if (Region *r = find_region(host, &base)) {
        if (base == (uintptr_t)host && r->bytes == bytes) return MSOD_OK;  // already present
        return fail(MSOD_ERR_ARG, "data_create: %p overlaps a present array", host);
    }
    auto it = g.present.lower_bound((uintptr_t)host);
    if (it != g.present.end() && it->first < (uintptr_t)host + bytes)
        return fail(MSOD_ERR_ARG, "data_create: %p+%zu overlaps a present array", host, bytes);
    void *dev = nullptr;
    HIPCHK(hipMalloc(&dev, bytes), "hipMalloc");
    hipError_t e = hipMemsetAsync(dev, 0, bytes, g.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g.stream);
    if (e != hipSuccess) {
        (void)hipFree(dev);
        return hipfail(e, "data_create");
    }
    g.present[(uintptr_t)host] = Region{bytes, dev, g.next_version++};
    return MSOD_OK;
}

int msod_data_copyin(const void *host, size_t bytes)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    NEED_INIT();
    uintptr_t base;
    Region *r0 = host ? find_region(host, &base) : nullptr;
    if (r0 && base == (uintptr_t)host && r0->bytes == bytes) return MSOD_OK;  // present: no copy
    int st = msod_data_create(host, bytes);
    if (st) return st;
    Region &r = g.present[(uintptr_t)host];
    HIPCHK(hipMemcpy(r.dev, host, bytes, hipMemcpyHostToDevice), "data_copyin");
    return MSOD_OK;
}

int msod_data_update_device(const void *host, size_t bytes)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    NEED_INIT();
    uintptr_t base;
    Region *r = host ? find_region(host, &base) : nullptr;
    if (!r) return fail(MSOD_ERR_NOT_PRESENT, "update_device: %p not present", host);
    if (!r->owned)
        return fail(MSOD_ERR_ARG, "update_device: %p is a registered device array (no host copy); "
                                  "call msod_data_invalidate after changing it", host);
    const size_t off = (uintptr_t)host - base;
    if (off + bytes > r->bytes) return fail(MSOD_ERR_ARG, "update_device: range exceeds the array");
    HIPCHK(hipStreamSynchronize(g.stream), "update_device");
    HIPCHK(hipMemcpy((char *)r->dev + off, host, bytes, hipMemcpyHostToDevice), "update_device");
    contents_changed(host);
    return MSOD_OK;
}

int msod_data_update_host(void *host, size_t bytes)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    NEED_INIT();
    uintptr_t base;
    Region *r = host ? find_region(host, &base) : nullptr;
    if (!r) return fail(MSOD_ERR_NOT_PRESENT, "update_host: %p not present", host);
    if (!r->owned)
        return fail(MSOD_ERR_ARG, "update_host: %p is a registered device array (no host copy)", host);
    const size_t off = (uintptr_t)host - base;
    if (off + bytes > r->bytes) return fail(MSOD_ERR_ARG, "update_host: range exceeds the array");
    HIPCHK(hipStreamSynchronize(g.stream), "update_host");
    HIPCHK(hipMemcpy(host, (const char *)r->dev + off, bytes, hipMemcpyDeviceToHost), "update_host");
    return MSOD_OK;
}

int msod_data_delete(const void *host)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    NEED_INIT();
    auto it = g.present.find((uintptr_t)host);
    if (it == g.present.end()) return fail(MSOD_ERR_NOT_PRESENT, "delete: %p not present", host);
    HIPCHK(hipStreamSynchronize(g.stream), "delete");
    const uintptr_t lo = (uintptr_t)it->second.dev, hi = lo + it->second.bytes;
    drop_derived_in(lo, hi);
    if (it->second.owned) HIPCHK(hipFree(it->second.dev), "hipFree");
    g.present.erase(it);
    return MSOD_OK;
}

int msod_data_register(const void *dev, size_t bytes)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    NEED_INIT();
    if (!dev || bytes == 0) return fail(MSOD_ERR_ARG, "data_register: NULL or empty");
    uintptr_t base;
    if (Region *r = find_region(dev, &base)) {
        if (base == (uintptr_t)dev && r->bytes == bytes && !r->owned) return MSOD_OK;  // already registered
        return fail(MSOD_ERR_ARG, "data_register: %p overlaps an array of the data region", dev);
    }
    if (find_region_dev(dev))
        return fail(MSOD_ERR_ARG, "data_register: %p is the device copy of a present array", dev);
    auto it = g.present.lower_bound((uintptr_t)dev);
    if (it != g.present.end() && it->first < (uintptr_t)dev + bytes)
        return fail(MSOD_ERR_ARG, "data_register: %p+%zu overlaps an array of the data region", dev, bytes);
    if (!is_device_pointer(dev)) return fail(MSOD_ERR_ARG, "data_register: %p is not device memory", dev);
    g.present[(uintptr_t)dev] = Region{bytes, const_cast<void *>(dev), g.next_version++, false};
    return MSOD_OK;
}

int msod_data_present(const void *host)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    uintptr_t base;
    return (g.init && host && find_region(host, &base)) ? 1 : 0;
}

void *msod_data_deviceptr(const void *host)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    uintptr_t base;
    Region *r = (g.init && host) ? find_region(host, &base) : nullptr;
    return r ? (char *)r->dev + ((uintptr_t)host - base) : nullptr;
}

int msod_data_invalidate(const void *ptr)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    NEED_INIT();
    HIPCHK(hipStreamSynchronize(g.stream), "invalidate");
    contents_changed(ptr);
    return MSOD_OK;
}

int msod_set_async(int async)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    g.async = async != 0;
    return MSOD_OK;
}

int msod_synchronize(void)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    NEED_INIT();
    HIPCHK(hipStreamSynchronize(g.stream), "synchronize");
    return MSOD_OK;
}

int msod_timing_enable(int enable)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    g.timing = enable != 0;
    return MSOD_OK;
}

int msod_timing_get(int kernel, double *total_ms, int64_t *launches)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    NEED_INIT();
    if (kernel < 0 || kernel >= kNumTimers) return fail(MSOD_ERR_ARG, "timer %d", kernel);
    collect_timing();
    if (total_ms) *total_ms = g.ms[kernel];
    if (launches) *launches = g.launches[kernel];
    return MSOD_OK;
}

int msod_timing_reset(void)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    NEED_INIT();
    collect_timing();
    for (int i = 0; i < kNumTimers; ++i) {
        g.ms[i] = 0.0;
        g.launches[i] = 0;
    }
    return MSOD_OK;
}

int msod_set_assembly(int mode)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    if (mode != MSOD_ASSEMBLY_COLORED && mode != MSOD_ASSEMBLY_ATOMIC)
        return fail(MSOD_ERR_ARG, "assembly mode %d", mode);
    g.assembly = mode == MSOD_ASSEMBLY_ATOMIC ? kAtomic : kColored;
    return MSOD_OK;
}

extern "C++" {
// Device copy of a small host vector, freed when the holder goes out of scope.
struct DevTemp {
    void *p = nullptr;
    ~DevTemp()
    {
        if (p) (void)hipFree(p);
    }
};
template <class V>
int dev_temp(const std::vector<V> &h, DevTemp &d, const char *what)
{
    HIPCHK(hipMalloc(&d.p, sizeof(V) * std::max<size_t>(1, h.size())), what);
    if (!h.empty()) HIPCHK(hipMemcpy(d.p, h.data(), sizeof(V) * h.size(), hipMemcpyHostToDevice), what);
    return MSOD_OK;
}
}  // extern "C++"

int msod_elem_jacobian(int porder, int nelem, int npoin, const int32_t *connec, const double *coord,
                       const double *dNgp, const double *wgp, double *gpvol, double *He, int64_t *nbad)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    MsodRange rg_("msod_elem_jacobian");
    NEED_INIT();
    int st = check_sizes(porder, nelem, npoin);
    if (st) return st;
    const int64_t nn = (int64_t)(porder + 1) * (porder + 1) * (porder + 1), NE = nelem, NP = npoin;
    const void *dc, *dx, *dv, *dh;
    uint64_t vx;
    if ((st = resolve(connec, sizeof(int32_t) * NE * nn, "connec", &dc, &vx))) return st;
    if ((st = resolve(coord, sizeof(double) * NP * 3, "coord", &dx, &vx))) return st;
    if ((st = resolve(gpvol, sizeof(double) * NE * nn, "gpvol", &dv, &vx))) return st;
    if ((st = resolve(He, sizeof(double) * NE * nn * 9, "He", &dh, &vx))) return st;
    // nonzero pattern of dNgp(3,nnode,ngaus), read on the host
    std::vector<char> t0, t1;
    const char *pd, *pw;
    if ((st = host_view(dNgp, sizeof(double) * 3 * nn * nn, t0, &pd, "dNgp"))) return st;
    if ((st = host_view(wgp, sizeof(double) * nn, t1, &pw, "wgp"))) return st;
    const double *D = (const double *)pd;
    std::vector<uint16_t> nzn(nn * 3, 0);
    int mz = 1;
    for (int64_t gg = 0; gg < nn; ++gg)
        for (int d = 0; d < 3; ++d) {
            int c = 0;
            for (int64_t a = 0; a < nn; ++a) c += D[d + 3 * (a + nn * gg)] != 0.0;
            nzn[gg * 3 + d] = (uint16_t)c;
            mz = std::max(mz, c);
        }
    std::vector<uint16_t> nza(nn * 3 * mz, 0);
    std::vector<double> nzw(nn * 3 * mz, 0.0);
    for (int64_t gg = 0; gg < nn; ++gg)
        for (int d = 0; d < 3; ++d) {
            int k = 0;
            for (int64_t a = 0; a < nn; ++a) {
                const double w = D[d + 3 * (a + nn * gg)];
                if (w != 0.0) {
                    nza[(gg * 3 + d) * mz + k] = (uint16_t)a;
                    nzw[(gg * 3 + d) * mz + k] = w;
                    ++k;
                }
            }
        }
    std::vector<double> w(pw ? (const double *)pw : nullptr, pw ? (const double *)pw + nn : nullptr);
    std::vector<unsigned long long> counts(2, 0ull);
    DevTemp da, dw, dn, dg, dk;
    if ((st = dev_temp(nza, da, "geometry tables"))) return st;
    if ((st = dev_temp(nzw, dw, "geometry tables"))) return st;
    if ((st = dev_temp(nzn, dn, "geometry tables"))) return st;
    if ((st = dev_temp(w, dg, "geometry tables"))) return st;
    if ((st = dev_temp(counts, dk, "geometry counters"))) return st;
    GeomArgs a;
    a.connec = (const int32_t *)dc;
    a.coord = (const double *)dx;
    a.wgp = (const double *)dg.p;
    a.nza = (const uint16_t *)da.p;
    a.nzw = (const double *)dw.p;
    a.nzn = (const uint16_t *)dn.p;
    a.nelem = NE;
    a.npoin = NP;
    a.mz = mz;
    a.He = (double *)dh;
    a.gpvol = (double *)dv;
    a.counts = (unsigned long long *)dk.p;
    TimedLaunch tk(9);
    HIPCHK(launch_elem_jacobian(porder, a, g.stream), "geometry kernel launch");
    tk.done();
    // synchronous: the temporaries are freed on return and the counts read
    HIPCHK(hipMemcpyAsync(counts.data(), dk.p, sizeof(unsigned long long) * 2, hipMemcpyDeviceToHost,
                          g.stream),
           "geometry counters");
    HIPCHK(hipStreamSynchronize(g.stream), "geometry kernel");
    // layouts derived from He and gpvol are stale now
    for (const void *ptr : {(const void *)He, (const void *)gpvol}) {
        contents_changed(ptr);
    }
    if (counts[0])
        return fail(MSOD_ERR_MESH, "connec holds %llu node numbers outside 1..npoin (npoin=%d)",
                    counts[0], npoin);
    if (nbad) *nbad = (int64_t)counts[1];
    return finish_call("geometry kernel");
}

int msod_box_mesh(int nx, int ny, int nz, int porder, double jitter, int64_t seed, const double *xgp,
                  const int32_t *AtoI, const int32_t *AtoJ, const int32_t *AtoK, int32_t *connec,
                  double *coord)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    MsodRange rg_("msod_box_mesh");
    NEED_INIT();
    if (porder < 1 || porder > kGenMaxP)
        return fail(MSOD_ERR_PORDER, "porder=%d: the generator supports 1..%d", porder, kGenMaxP);
    if (nx < 1 || ny < 1 || nz < 1 || !(jitter >= 0.0 && jitter < 1.0))
        return fail(MSOD_ERR_ARG, "box %d x %d x %d, jitter %g", nx, ny, nz, jitter);
    const int q = porder + 1, nn = q * q * q;
    BoxArgs b{};
    b.nx = nx;
    b.ny = ny;
    b.nz = nz;
    b.p = porder;
    b.nn = nn;
    b.nelem = (int64_t)nx * ny * nz;
    b.n1x = (int64_t)nx * porder + 1;
    b.n1y = (int64_t)ny * porder + 1;
    b.npoin = b.n1x * b.n1y * ((int64_t)nz * porder + 1);
    if (b.npoin >= (int64_t)1 << 31 || b.nelem * nn >= (int64_t)1 << 31)
        return fail(MSOD_ERR_ARG, "mesh too large for the int32 index format of the reference");
    b.h[0] = 1.0 / nx;
    b.h[1] = 1.0 / ny;
    b.h[2] = 1.0 / nz;
    b.jitter = jitter;
    b.seed = (uint64_t)seed;
    // lattice positions from the SOD 1-D indices (1: xi=-1, 2: xi=+1, 3..q:
    // interior ascending) and the 1-D points from xgp(nnode,3)
    std::vector<char> t0, t1, t2, t3;
    const char *px, *pi, *pj, *pk;
    int st;
    if ((st = host_view(xgp, sizeof(double) * nn * 3, t0, &px, "xgp"))) return st;
    if ((st = host_view(AtoI, sizeof(int32_t) * nn, t1, &pi, "gmshAtoI"))) return st;
    if ((st = host_view(AtoJ, sizeof(int32_t) * nn, t2, &pj, "gmshAtoJ"))) return st;
    if ((st = host_view(AtoK, sizeof(int32_t) * nn, t3, &pk, "gmshAtoK"))) return st;
    const int32_t *A[3] = {(const int32_t *)pi, (const int32_t *)pj, (const int32_t *)pk};
    const double *X = (const double *)px;
    std::vector<int8_t> pos(3 * nn);
    std::vector<int> seen(q, 0);
    for (int a = 0; a < nn; ++a)
        for (int d = 0; d < 3; ++d) {
            const int i = A[d][a];
            if (i < 1 || i > q) return fail(MSOD_ERR_TABLE, "gmshAto%c(%d) = %d", "IJK"[d], a + 1, i);
            const int ps = i == 1 ? 0 : (i == 2 ? porder : i - 2);
            pos[3 * a + d] = (int8_t)ps;
            const double x = X[a + (int64_t)nn * d];
            if (seen[ps] && b.xi[ps] != x) return fail(MSOD_ERR_TABLE, "xgp inconsistent with gmshAtoI/J/K");
            b.xi[ps] = x;
            seen[ps] = 1;
        }
    for (int ps = 1; ps < porder; ++ps)
        b.hloc[ps] = std::min(b.xi[ps + 1] - b.xi[ps], b.xi[ps] - b.xi[ps - 1]);
    const void *dc, *dx;
    uint64_t vx;
    if ((st = resolve(connec, sizeof(int32_t) * b.nelem * nn, "connec", &dc, &vx))) return st;
    if ((st = resolve(coord, sizeof(double) * b.npoin * 3, "coord", &dx, &vx))) return st;
    DevTemp dpos;
    if ((st = dev_temp(pos, dpos, "generator tables"))) return st;
    b.pos = (const int8_t *)dpos.p;
    b.connec = (int32_t *)dc;
    b.coord = (double *)dx;
    TimedLaunch tk(10);
    HIPCHK(launch_box_mesh(b, g.stream), "box mesh launch");
    tk.done();
    HIPCHK(hipStreamSynchronize(g.stream), "box mesh");  // the table is freed on return
    for (const void *ptr : {(const void *)connec, (const void *)coord}) {
        contents_changed(ptr);
    }
    return finish_call("box mesh");
}

int msod_flow_fields(int64_t npoin, const double *coord, double gamma, double Rgas, int64_t seed_rho,
                     double rho_jump, double *u, double *q, double *rho, double *pr, double *E,
                     double *Tem, double *mu_fluid)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    MsodRange rg_("msod_flow_fields");
    NEED_INIT();
    if (npoin < 1 || !(gamma > 1.0) || !(Rgas > 0.0))
        return fail(MSOD_ERR_ARG, "npoin=%lld gamma=%g Rgas=%g", (long long)npoin, gamma, Rgas);
    FlowArgs f{};
    f.npoin = npoin;
    f.gamma = gamma;
    f.Rgas = Rgas;
    f.rho_jump = rho_jump;
    f.seed_rho = (uint64_t)seed_rho;
    const size_t n8 = sizeof(double) * (size_t)npoin;
    const void *d[9];
    uint64_t vx;
    int st;
    const void *in[9] = {coord, u, q, rho, pr, E, Tem, mu_fluid, nullptr};
    const size_t sz[8] = {3 * n8, 3 * n8, 3 * n8, n8, n8, n8, n8, n8};
    const char *nm[8] = {"coord", "u", "q", "rho", "pr", "E", "Tem", "mu_fluid"};
    for (int k = 0; k < 8; ++k)
        if ((st = resolve(in[k], sz[k], nm[k], &d[k], &vx))) return st;
    f.coord = (const double *)d[0];
    f.u = (double *)d[1];
    f.q = (double *)d[2];
    f.rho = (double *)d[3];
    f.pr = (double *)d[4];
    f.E = (double *)d[5];
    f.Tem = (double *)d[6];
    f.mu_fluid = (double *)d[7];
    TimedLaunch tk(10);
    HIPCHK(launch_flow_fields(f, g.stream), "flow fields launch");
    tk.done();
    for (int k = 1; k < 8; ++k) {
        uintptr_t base;
        (void)base;
        bump_version(in[k]);
    }
    return finish_call("flow fields");
}

int msod_viscosity_fields(int64_t nelem, int ngaus, int64_t seed_mue, int64_t seed_musgs, double mue_max,
                          double musgs_max, double *mu_e, double *mu_sgs)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    MsodRange rg_("msod_viscosity_fields");
    NEED_INIT();
    if (nelem < 1 || ngaus < 1) return fail(MSOD_ERR_ARG, "nelem=%lld ngaus=%d", (long long)nelem, ngaus);
    ViscArgs v{};
    v.nelem = nelem;
    v.ngaus = ngaus;
    v.seed_mue = (uint64_t)seed_mue;
    v.seed_musgs = (uint64_t)seed_musgs;
    v.mue_max = mue_max;
    v.musgs_max = musgs_max;
    const void *d0, *d1;
    uint64_t vx;
    int st;
    if ((st = resolve(mu_e, sizeof(double) * nelem * ngaus, "mu_e", &d0, &vx))) return st;
    if ((st = resolve(mu_sgs, sizeof(double) * nelem * ngaus, "mu_sgs", &d1, &vx))) return st;
    v.mu_e = (double *)d0;
    v.mu_sgs = (double *)d1;
    TimedLaunch tk(10);
    HIPCHK(launch_viscosity_fields(v, g.stream), "viscosity fields launch");
    tk.done();
    for (const void *ptr : {(const void *)mu_e, (const void *)mu_sgs}) {
        uintptr_t base;
        (void)base;
        bump_version(ptr);
    }
    return finish_call("viscosity fields");
}

// ---- linear tetrahedra ------------------------------------------------------
extern "C++" {
// Ngp(4,4), dNgp(3,4,4), wgp(4) of the TET04 element, read on the host
static int tet_tables(const double *Ngp, const double *dNgp, const double *wgp, TetTables &t)
{
    std::vector<char> t0, t1, t2;
    const char *pn = nullptr, *pd = nullptr, *pw = nullptr;
    int st;
    if (Ngp && (st = host_view(Ngp, sizeof t.Ngp, t0, &pn, "Ngp"))) return st;
    if (dNgp && (st = host_view(dNgp, sizeof t.dNgp, t1, &pd, "dNgp"))) return st;
    if (wgp && (st = host_view(wgp, sizeof t.wgp, t2, &pw, "wgp"))) return st;
    std::memset(&t, 0, sizeof t);
    if (pn) std::memcpy(t.Ngp, pn, sizeof t.Ngp);
    if (pd) std::memcpy(t.dNgp, pd, sizeof t.dNgp);
    if (pw) std::memcpy(t.wgp, pw, sizeof t.wgp);
    for (double v : t.Ngp)
        if (!std::isfinite(v)) return fail(MSOD_ERR_TABLE, "Ngp is not finite");
    for (double v : t.dNgp)
        if (!std::isfinite(v)) return fail(MSOD_ERR_TABLE, "dNgp is not finite");
    return MSOD_OK;
}

// connec(nelem,4) analysed on the device once per array version: entries
// outside 1..npoin are an error (the kernels clamp in any case, so they never
// read out of bounds); a mesh whose nodes each belong to at most one element
// node (the reference's TET04 mesh) gets plain stores and the list of the
// nodes it leaves out.
static int check_tet_connec(const void *dc, uint64_t ver, int64_t nelem, int64_t npoin,
                            const CheckedMesh **out)
{
    const int64_t n = 4 * nelem;
    if (ver == 0) drop_unversioned(dc);
    for (const CheckedMesh &c : g.checked)
        if (c.src == dc && c.ver == ver && c.n == n && c.npoin == npoin) {
            *out = &c;
            return MSOD_OK;
        }
    std::vector<unsigned long long> stats(4, 0ull);
    DevTemp dk, drefs;
    int st;
    if ((st = dev_temp(stats, dk, "mesh analysis"))) return st;
    unsigned long long *ds = (unsigned long long *)dk.p;
    const int64_t nwords = (npoin + 1) / 2;  // the counters, zeroed as doubles
    HIPCHK(hipMalloc(&drefs.p, sizeof(double) * nwords), "mesh analysis buffer");
    HIPCHK(launch_zero_doubles((double *)drefs.p, nwords, g.stream), "mesh analysis");
    HIPCHK(launch_count_refs((const int32_t *)dc, n, npoin, (uint32_t *)drefs.p, ds, g.stream),
           "mesh analysis launch");
    HIPCHK(launch_ref_summary((const uint32_t *)drefs.p, npoin, ds, g.stream), "mesh analysis launch");
    HIPCHK(hipMemcpyAsync(stats.data(), ds, sizeof(unsigned long long) * 3, hipMemcpyDeviceToHost, g.stream),
           "mesh analysis");
    HIPCHK(hipStreamSynchronize(g.stream), "mesh analysis");
    if (stats[0])
        return fail(MSOD_ERR_MESH, "connec holds %llu node numbers outside 1..npoin (npoin=%lld)", stats[0],
                    (long long)npoin);
    CheckedMesh c{dc, ver, n, npoin, stats[1] == 0, nullptr, 0};
    if (c.disjoint && stats[2] > 0) {
        c.nunref = (int64_t)stats[2];
        HIPCHK(hipMalloc((void **)&c.unref, sizeof(uint32_t) * c.nunref), "unreferenced-node list");
        hipError_t e = launch_list_unreferenced((const uint32_t *)drefs.p, npoin, c.unref, ds + 3, g.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(g.stream);
        if (e != hipSuccess) {
            c.release();
            return hipfail(e, "unreferenced-node list");
        }
    }
    if (g.checked.size() >= kMaxDerived) {
        g.checked.front().release();
        g.checked.erase(g.checked.begin());
    }
    g.checked.push_back(c);
    *out = &g.checked.back();
    return MSOD_OK;
}

// Outputs before a TET04 kernel: zeroed where the kernel will not store.
static int tet_zero(const CheckedMesh &cm, double *Rmass, double *Rmom, double *Rener, int64_t npoin)
{
    TimedLaunch tz(0);
    if (!cm.disjoint)
        HIPCHK(launch_zero_outputs(Rmass, Rmom, Rener, npoin, g.stream), "zero-fill launch");
    else if (cm.nunref > 0)
        HIPCHK(launch_zero_nodes(Rmass, Rmom, Rener, npoin, cm.unref, cm.nunref, g.stream), "zero-fill launch");
    tz.done();
    return MSOD_OK;
}

static int check_tet_sizes(int nelem, int npoin)
{
    if (nelem <= 0 || npoin <= 0) return fail(MSOD_ERR_ARG, "nelem=%d npoin=%d", nelem, npoin);
    return MSOD_OK;
}
}  // extern "C++"

int msod_tet_jacobian(int nelem, int npoin, const int32_t *connec, const double *coord,
                      const double *dNgp, const double *wgp, double *gpvol, double *He, int64_t *nbad)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    MsodRange rg_("msod_tet_jacobian");
    NEED_INIT();
    int st = check_tet_sizes(nelem, npoin);
    if (st) return st;
    const int64_t NE = nelem, NP = npoin;
    const void *dc, *dx, *dv, *dh;
    uint64_t vx;
    if ((st = resolve(connec, sizeof(int32_t) * NE * 4, "connec", &dc, &vx))) return st;
    if ((st = resolve(coord, sizeof(double) * NP * 3, "coord", &dx, &vx))) return st;
    if ((st = resolve(gpvol, sizeof(double) * NE * 4, "gpvol", &dv, &vx))) return st;
    if ((st = resolve(He, sizeof(double) * NE * 36, "He", &dh, &vx))) return st;
    TetJacobianArgs a;
    if (!dNgp || !wgp) return fail(MSOD_ERR_ARG, "dNgp and wgp are required");
    if ((st = tet_tables(nullptr, dNgp, wgp, a.t))) return st;
    std::vector<unsigned long long> counts(2, 0ull);
    DevTemp dk;
    if ((st = dev_temp(counts, dk, "geometry counters"))) return st;
    a.nelem = NE;
    a.npoin = NP;
    a.connec = (const int32_t *)dc;
    a.coord = (const double *)dx;
    a.He = (double *)dh;
    a.gpvol = (double *)dv;
    a.counts = (unsigned long long *)dk.p;
    TimedLaunch tk(13);
    HIPCHK(launch_tet_jacobian(a, g.stream), "TET geometry kernel launch");
    tk.done();
    HIPCHK(hipMemcpyAsync(counts.data(), dk.p, sizeof(unsigned long long) * 2, hipMemcpyDeviceToHost,
                          g.stream),
           "TET geometry counters");
    HIPCHK(hipStreamSynchronize(g.stream), "TET geometry kernel");
    for (const void *ptr : {(const void *)He, (const void *)gpvol}) {
        contents_changed(ptr);
    }
    if (counts[0])
        return fail(MSOD_ERR_MESH, "connec holds %llu node numbers outside 1..npoin (npoin=%d)",
                    counts[0], npoin);
    if (nbad) *nbad = (int64_t)counts[1];
    return finish_call("TET geometry kernel");
}

int msod_fem_convec(int nelem, int npoin, const int32_t *connec, const double *Ngp, const double *dNgp,
                    const double *He, const double *gpvol, const double *u, const double *q,
                    const double *rho, const double *pr, const double *E, double *Rmass, double *Rmom,
                    double *Rener)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    MsodRange rg_("msod_fem_convec");
    NEED_INIT();
    int st = check_tet_sizes(nelem, npoin);
    if (st) return st;
    const int64_t NE = nelem, NP = npoin;
    TetConvecArgs a;
    if (!Ngp || !dNgp) return fail(MSOD_ERR_ARG, "Ngp and dNgp are required");
    if ((st = tet_tables(Ngp, dNgp, nullptr, a.t))) return st;
    const void *dc, *dh, *dv, *du, *dq, *dr, *dp, *dE, *dRm, *dRmo, *dRe;
    uint64_t vc, vx;
    if ((st = resolve(connec, sizeof(int32_t) * NE * 4, "connec", &dc, &vc))) return st;
    if ((st = resolve(He, sizeof(double) * NE * 36, "He", &dh, &vx))) return st;
    if ((st = resolve(gpvol, sizeof(double) * NE * 4, "gpvol", &dv, &vx))) return st;
    if ((st = resolve(u, sizeof(double) * NP * 3, "u", &du, &vx))) return st;
    if ((st = resolve(q, sizeof(double) * NP * 3, "q", &dq, &vx))) return st;
    if ((st = resolve(rho, sizeof(double) * NP, "rho", &dr, &vx))) return st;
    if ((st = resolve(pr, sizeof(double) * NP, "pr", &dp, &vx))) return st;
    if ((st = resolve(E, sizeof(double) * NP, "E", &dE, &vx))) return st;
    if ((st = resolve(Rmass, sizeof(double) * NP, "Rmass", &dRm, &vx))) return st;
    if ((st = resolve(Rmom, sizeof(double) * NP * 3, "Rmom", &dRmo, &vx))) return st;
    if ((st = resolve(Rener, sizeof(double) * NP, "Rener", &dRe, &vx))) return st;
    const CheckedMesh *cm;
    if ((st = check_tet_connec(dc, vc, NE, NP, &cm))) return st;
    a.nelem = NE;
    a.npoin = NP;
    a.connec = (const int32_t *)dc;
    a.He = (const double *)dh;
    a.gpvol = (const double *)dv;
    a.u = (const double *)du;
    a.q = (const double *)dq;
    a.rho = (const double *)dr;
    a.pr = (const double *)dp;
    a.E = (const double *)dE;
    a.Rmass = (double *)dRm;
    a.Rmom = (double *)dRmo;
    a.Rener = (double *)dRe;
    a.store = cm->disjoint;
    if ((st = tet_zero(*cm, a.Rmass, a.Rmom, a.Rener, NP))) return st;
    TimedLaunch tk(11);
    HIPCHK(launch_fem_convec(a, g.stream), "TET convection kernel launch");
    tk.done();
    return finish_call("TET convection kernel");
}

int msod_fem_diffu(int nelem, int npoin, const int32_t *connec, const double *Ngp, const double *dNgp,
                   const double *He, const double *gpvol, double Cp, double Pr, const double *rho,
                   const double *u, const double *Tem, const double *mu_fluid, const double *mu_e,
                   const double *mu_sgs, double *Rmass, double *Rmom, double *Rener)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    MsodRange rg_("msod_fem_diffu");
    NEED_INIT();
    int st = check_tet_sizes(nelem, npoin);
    if (st) return st;
    if (!(Pr != 0.0) || !std::isfinite(Cp) || !std::isfinite(Pr))
        return fail(MSOD_ERR_ARG, "Cp=%g Pr=%g", Cp, Pr);
    const int64_t NE = nelem, NP = npoin;
    TetDiffuArgs a;
    if (!Ngp || !dNgp) return fail(MSOD_ERR_ARG, "Ngp and dNgp are required");
    if ((st = tet_tables(Ngp, dNgp, nullptr, a.t))) return st;
    const void *dc, *dh, *dv, *dr, *du, *dT, *dmf, *dme, *dms, *dRm, *dRmo, *dRe;
    uint64_t vc, vx;
    if ((st = resolve(connec, sizeof(int32_t) * NE * 4, "connec", &dc, &vc))) return st;
    if ((st = resolve(He, sizeof(double) * NE * 36, "He", &dh, &vx))) return st;
    if ((st = resolve(gpvol, sizeof(double) * NE * 4, "gpvol", &dv, &vx))) return st;
    if ((st = resolve(rho, sizeof(double) * NP, "rho", &dr, &vx))) return st;
    if ((st = resolve(u, sizeof(double) * NP * 3, "u", &du, &vx))) return st;
    if ((st = resolve(Tem, sizeof(double) * NP, "Tem", &dT, &vx))) return st;
    if ((st = resolve(mu_fluid, sizeof(double) * NP, "mu_fluid", &dmf, &vx))) return st;
    if ((st = resolve(mu_e, sizeof(double) * NE * 4, "mu_e", &dme, &vx))) return st;
    if ((st = resolve(mu_sgs, sizeof(double) * NE * 4, "mu_sgs", &dms, &vx))) return st;
    if ((st = resolve(Rmass, sizeof(double) * NP, "Rmass", &dRm, &vx))) return st;
    if ((st = resolve(Rmom, sizeof(double) * NP * 3, "Rmom", &dRmo, &vx))) return st;
    if ((st = resolve(Rener, sizeof(double) * NP, "Rener", &dRe, &vx))) return st;
    const CheckedMesh *cm;
    if ((st = check_tet_connec(dc, vc, NE, NP, &cm))) return st;
    a.nelem = NE;
    a.npoin = NP;
    a.connec = (const int32_t *)dc;
    a.He = (const double *)dh;
    a.gpvol = (const double *)dv;
    a.rho = (const double *)dr;
    a.u = (const double *)du;
    a.Tem = (const double *)dT;
    a.mu_fluid = (const double *)dmf;
    a.mu_e = (const double *)dme;
    a.mu_sgs = (const double *)dms;
    a.Cp = Cp;
    a.Pr = Pr;
    a.c_rho = g.c_rho;
    a.c_ener = g.c_ener;
    a.Rmass = (double *)dRm;
    a.Rmom = (double *)dRmo;
    a.Rener = (double *)dRe;
    a.store = cm->disjoint;
    if ((st = tet_zero(*cm, a.Rmass, a.Rmom, a.Rener, NP))) return st;
    TimedLaunch tk(12);
    HIPCHK(launch_fem_diffu(a, g.stream), "TET diffusion kernel launch");
    tk.done();
    return finish_call("TET diffusion kernel");
}

int msod_set_diffusion_coefficients(double c_rho, double c_ener)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    g.c_rho = c_rho;
    g.c_ener = c_ener;
    return MSOD_OK;
}

int msod_full_convec_ijk(int porder, int nelem, int npoin, const int32_t *connec, const double *Ngp,
                         const double *dNgp, const double *He, const double *gpvol,
                         const double *dlxigp_ip, const double *xgp, const int32_t *invAtoIJK,
                         const int32_t *gmshAtoI, const int32_t *gmshAtoJ, const int32_t *gmshAtoK,
                         const double *u, const double *q, const double *rho, const double *pr,
                         const double *E, double *Rmass, double *Rmom, double *Rener)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    MsodRange rg_("msod_full_convec_ijk");
    (void)Ngp; (void)dNgp; (void)xgp;
    if (!pr) return fail(MSOD_ERR_ARG, "pr is NULL");
    return convec_common<double, double>(porder, nelem, npoin, connec, He, gpvol, dlxigp_ip, invAtoIJK,
                                 gmshAtoI, gmshAtoJ, gmshAtoK, u, q, rho, pr, E, 0.0, Rmass, Rmom,
                                 Rener, 1);
}

int msod_full_convec_ijk_mixed(int porder, int nelem, int npoin, const int32_t *connec,
                               const float *He, const float *gpvol, const double *dlxigp_ip,
                               const int32_t *invAtoIJK, const int32_t *gmshAtoI,
                               const int32_t *gmshAtoJ, const int32_t *gmshAtoK, const float *u,
                               const float *q, const float *rho, const float *E, double gamma_gas,
                               double *Rmass, double *Rmom, double *Rener)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    MsodRange rg_("msod_full_convec_ijk_mixed");
    if (!(gamma_gas > 1.0)) return fail(MSOD_ERR_ARG, "gamma_gas=%g must exceed 1", gamma_gas);
    return convec_common<float, float>(porder, nelem, npoin, connec, He, gpvol, dlxigp_ip, invAtoIJK,
                                gmshAtoI, gmshAtoJ, gmshAtoK, u, q, rho, (const float *)nullptr, E, gamma_gas,
                                Rmass, Rmom, Rener, 3);
}

int msod_full_diffusion_ijk(int porder, int nelem, int npoin, const int32_t *connec,
                            const double *Ngp, const double *dNgp, const double *He,
                            const double *gpvol, const double *dlxigp_ip, const double *xgp,
                            const int32_t *invAtoIJK, const int32_t *gmshAtoI,
                            const int32_t *gmshAtoJ, const int32_t *gmshAtoK, double Cp, double Pr,
                            const double *rho, const double *u, const double *Tem,
                            const double *mu_fluid, const double *mu_e, const double *mu_sgs,
                            double *Rmass, double *Rmom, double *Rener)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    MsodRange rg_("msod_full_diffusion_ijk");
    (void)Ngp; (void)dNgp; (void)xgp;
    return diffu_common<double, double>(porder, nelem, npoin, connec, He, gpvol, dlxigp_ip, invAtoIJK,
                                gmshAtoI, gmshAtoJ, gmshAtoK, Cp, Pr, rho, u, Tem, mu_fluid, mu_e,
                                mu_sgs, Rmass, Rmom, Rener, 2);
}

int msod_full_diffusion_ijk_mixed(int porder, int nelem, int npoin, const int32_t *connec,
                                  const float *He, const float *gpvol, const double *dlxigp_ip,
                                  const int32_t *invAtoIJK, const int32_t *gmshAtoI,
                                  const int32_t *gmshAtoJ, const int32_t *gmshAtoK, double Cp,
                                  double Pr, const float *rho, const float *u, const float *Tem,
                                  const float *mu_fluid, const float *mu_e, const float *mu_sgs,
                                  double *Rmass, double *Rmom, double *Rener)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    MsodRange rg_("msod_full_diffusion_ijk_mixed");
    return diffu_common<float, float>(porder, nelem, npoin, connec, He, gpvol, dlxigp_ip, invAtoIJK,
                               gmshAtoI, gmshAtoJ, gmshAtoK, Cp, Pr, rho, u, Tem, mu_fluid, mu_e,
                               mu_sgs, Rmass, Rmom, Rener, 4);
}

int msod_full_convec_ijk_f32split(int porder, int nelem, int npoin, const int32_t *connec, const float *He,
                                  const float *gpvol, const double *dlxigp_ip, const int32_t *invAtoIJK,
                                  const int32_t *gmshAtoI, const int32_t *gmshAtoJ, const int32_t *gmshAtoK,
                                  const double *u, const float *q, const float *rho, const double *pr,
                                  const float *E, double *Rmass, double *Rmom, double *Rener)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    MsodRange rg_("msod_full_convec_ijk_f32split");
    if (!pr) return fail(MSOD_ERR_ARG, "pr is NULL");
    return convec_common<float, float, double, double, double>(
        porder, nelem, npoin, connec, He, gpvol, dlxigp_ip, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, u, q, rho,
        pr, E, 0.0, Rmass, Rmom, Rener, 18);
}

int msod_full_diffusion_ijk_f32split(int porder, int nelem, int npoin, const int32_t *connec, const float *He,
                                     const float *gpvol, const double *dlxigp_ip, const int32_t *invAtoIJK,
                                     const int32_t *gmshAtoI, const int32_t *gmshAtoJ, const int32_t *gmshAtoK,
                                     double Cp, double Pr, const float *rho, const double *u, const float *Tem,
                                     const float *mu_fluid, const float *mu_e, const float *mu_sgs,
                                     double *Rmass, double *Rmom, double *Rener)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    MsodRange rg_("msod_full_diffusion_ijk_f32split");
    return diffu_common<float, float, double, double>(porder, nelem, npoin, connec, He, gpvol, dlxigp_ip,
                                                      invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, Cp, Pr, rho, u,
                                                      Tem, mu_fluid, mu_e, mu_sgs, Rmass, Rmom, Rener, 19);
}

int msod_full_convec_ijk_r4(int porder, int nelem, int npoin, const int32_t *connec, const float *Ngp,
                            const float *dNgp, const float *He, const float *gpvol,
                            const float *dlxigp_ip, const float *xgp, const int32_t *invAtoIJK,
                            const int32_t *gmshAtoI, const int32_t *gmshAtoJ, const int32_t *gmshAtoK,
                            const float *u, const float *q, const float *rho, const float *pr,
                            const float *E, float *Rmass, float *Rmom, float *Rener)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    MsodRange rg_("msod_full_convec_ijk_r4");
    (void)Ngp; (void)dNgp; (void)xgp;
    if (!pr) return fail(MSOD_ERR_ARG, "pr is NULL");
    return convec_common<float, float, float>(porder, nelem, npoin, connec, He, gpvol, dlxigp_ip, invAtoIJK,
                                              gmshAtoI, gmshAtoJ, gmshAtoK, u, q, rho, pr, E, 0.0, Rmass,
                                              Rmom, Rener, 16);
}

int msod_full_diffusion_ijk_r4(int porder, int nelem, int npoin, const int32_t *connec, const float *Ngp,
                               const float *dNgp, const float *He, const float *gpvol,
                               const float *dlxigp_ip, const float *xgp, const int32_t *invAtoIJK,
                               const int32_t *gmshAtoI, const int32_t *gmshAtoJ, const int32_t *gmshAtoK,
                               float Cp, float Pr, const float *rho, const float *u, const float *Tem,
                               const float *mu_fluid, const float *mu_e, const float *mu_sgs,
                               float *Rmass, float *Rmom, float *Rener)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    MsodRange rg_("msod_full_diffusion_ijk_r4");
    (void)Ngp; (void)dNgp; (void)xgp;
    return diffu_common<float, float, float>(porder, nelem, npoin, connec, He, gpvol, dlxigp_ip, invAtoIJK,
                                             gmshAtoI, gmshAtoJ, gmshAtoK, (double)Cp, (double)Pr, rho, u, Tem,
                                             mu_fluid, mu_e, mu_sgs, Rmass, Rmom, Rener, 17);
}

int msod_full_convec_ijk_fp32(int porder, int nelem, int npoin, const int32_t *connec,
                              const float *He, const float *gpvol, const double *dlxigp_ip,
                              const int32_t *invAtoIJK, const int32_t *gmshAtoI,
                              const int32_t *gmshAtoJ, const int32_t *gmshAtoK, const float *u,
                              const float *q, const float *rho, const float *E, double gamma_gas,
                              double *Rmass, double *Rmom, double *Rener)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    MsodRange rg_("msod_full_convec_ijk_fp32");
    if (!(gamma_gas > 1.0)) return fail(MSOD_ERR_ARG, "gamma_gas=%g must exceed 1", gamma_gas);
    return convec_common<float, float>(porder, nelem, npoin, connec, He, gpvol, dlxigp_ip, invAtoIJK,
                                gmshAtoI, gmshAtoJ, gmshAtoK, u, q, rho, (const float *)nullptr, E, gamma_gas,
                                Rmass, Rmom, Rener, 14, true);
}

int msod_full_diffusion_ijk_fp32(int porder, int nelem, int npoin, const int32_t *connec,
                                 const float *He, const float *gpvol, const double *dlxigp_ip,
                                 const int32_t *invAtoIJK, const int32_t *gmshAtoI,
                                 const int32_t *gmshAtoJ, const int32_t *gmshAtoK, double Cp,
                                 double Pr, const float *rho, const float *u, const float *Tem,
                                 const float *mu_fluid, const float *mu_e, const float *mu_sgs,
                                 double *Rmass, double *Rmom, double *Rener)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    MsodRange rg_("msod_full_diffusion_ijk_fp32");
    return diffu_common<float, float>(porder, nelem, npoin, connec, He, gpvol, dlxigp_ip, invAtoIJK,
                               gmshAtoI, gmshAtoJ, gmshAtoK, Cp, Pr, rho, u, Tem, mu_fluid, mu_e,
                               mu_sgs, Rmass, Rmom, Rener, 15, false, nullptr, true);
}

int msod_full_convec_ijk_geo32(int porder, int nelem, int npoin, const int32_t *connec,
                               const float *He, const float *gpvol, const double *dlxigp_ip,
                               const int32_t *invAtoIJK, const int32_t *gmshAtoI,
                               const int32_t *gmshAtoJ, const int32_t *gmshAtoK, const double *u,
                               const double *q, const double *rho, const double *pr,
                               const double *E, double *Rmass, double *Rmom, double *Rener)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    MsodRange rg_("msod_full_convec_ijk_geo32");
    if (!pr) return fail(MSOD_ERR_ARG, "pr is NULL");
    return convec_common<float, double>(porder, nelem, npoin, connec, He, gpvol, dlxigp_ip, invAtoIJK,
                                        gmshAtoI, gmshAtoJ, gmshAtoK, u, q, rho, pr, E, 0.0, Rmass,
                                        Rmom, Rener, 7);
}

int msod_full_diffusion_ijk_geo32(int porder, int nelem, int npoin, const int32_t *connec,
                                  const float *He, const float *gpvol, const double *dlxigp_ip,
                                  const int32_t *invAtoIJK, const int32_t *gmshAtoI,
                                  const int32_t *gmshAtoJ, const int32_t *gmshAtoK, double Cp,
                                  double Pr, const double *rho, const double *u, const double *Tem,
                                  const double *mu_fluid, const float *mu_e, const float *mu_sgs,
                                  double *Rmass, double *Rmom, double *Rener)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    MsodRange rg_("msod_full_diffusion_ijk_geo32");
    return diffu_common<float, double>(porder, nelem, npoin, connec, He, gpvol, dlxigp_ip, invAtoIJK,
                                       gmshAtoI, gmshAtoJ, gmshAtoK, Cp, Pr, rho, u, Tem, mu_fluid, mu_e,
                                       mu_sgs, Rmass, Rmom, Rener, 8);
}

int msod_full_diffusion_ijk_closure(int porder, int nelem, int npoin, const int32_t *connec,
                                    const double *Ngp, const double *dNgp, const double *He,
                                    const double *gpvol, const double *dlxigp_ip,
                                    const double *xgp, const int32_t *invAtoIJK,
                                    const int32_t *gmshAtoI, const int32_t *gmshAtoJ,
                                    const int32_t *gmshAtoK, double Cp, double Pr,
                                    const double *rho, const double *u, const double *Tem,
                                    const double *mu_fluid, const double *mu_e,
                                    const float *mlp_params, int *used_model, double *Rmass,
                                    double *Rmom, double *Rener)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    MsodRange rg_("msod_full_diffusion_ijk_closure");
    (void)Ngp; (void)dNgp; (void)xgp;
    NEED_INIT();
    if (used_model) *used_model = 0;
    const float *dmlp = nullptr;
    bool usable = false;
    int st = closure_weights(mlp_params, &dmlp, &usable);
    if (st) return st;
    if (used_model) *used_model = usable ? 1 : 0;
    // unusable model: mu_sgs = 0 (the plain kernel with no mu_sgs array)
    return diffu_common<double, double>(porder, nelem, npoin, connec, He, gpvol, dlxigp_ip, invAtoIJK,
                                gmshAtoI, gmshAtoJ, gmshAtoK, Cp, Pr, rho, u, Tem, mu_fluid, mu_e,
                                nullptr, Rmass, Rmom, Rener, 5, true, usable ? dmlp : nullptr);
}

int msod_closure_features(int porder, int nelem, int npoin, const int32_t *connec,
                          const double *He, const double *gpvol, const double *dlxigp_ip,
                          const int32_t *invAtoIJK, const int32_t *gmshAtoI,
                          const int32_t *gmshAtoJ, const int32_t *gmshAtoK, const double *rho,
                          const double *u, float *features)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    MsodRange rg_("msod_closure_features");
    ClosureElemArgs a;
    int st = closure_elem_common(porder, nelem, npoin, connec, He, gpvol, dlxigp_ip, invAtoIJK,
                                 gmshAtoI, gmshAtoJ, gmshAtoK, rho, u, a);
    if (st) return st;
    const int64_t nn = (int64_t)(porder + 1) * (porder + 1) * (porder + 1);
    const void *df;
    uint64_t vx;
    if ((st = resolve(features, sizeof(float) * kMlpIn * nn * nelem, "features", &df, &vx))) return st;
    a.feat = (float *)df;
    TimedLaunch tk(6);
    HIPCHK(launch_closure_features(porder, a, g.stream), "closure features launch");
    tk.done();
    return finish_call("closure features");
}

int msod_closure_predict(int porder, int nelem, const int32_t *invAtoIJK, const float *mlp_params,
                         const float *features, double *mu_sgs)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    MsodRange rg_("msod_closure_predict");
    NEED_INIT();
    int st = check_sizes(porder, nelem, 1);
    if (st) return st;
    const int nn = (porder + 1) * (porder + 1) * (porder + 1);
    const int32_t *dl2g;
    if ((st = get_l2g(porder, invAtoIJK, &dl2g))) return st;
    const float *dmlp;
    bool usable;
    if ((st = closure_weights(mlp_params, &dmlp, &usable))) return st;
    if (!usable) return fail(MSOD_ERR_MODEL, "mlp_params missing, not finite or with a negative scale");
    const void *df, *dm;
    uint64_t vx;
    if ((st = resolve(features, sizeof(float) * kMlpIn * nn * (int64_t)nelem, "features", &df, &vx))) return st;
    if ((st = resolve(mu_sgs, sizeof(double) * nn * (int64_t)nelem, "mu_sgs", &dm, &vx))) return st;
    TimedLaunch tk(6);
    HIPCHK(launch_mlp_predict(dmlp, (const float *)df, nelem, nn, dl2g, (double *)dm, g.stream),
           "closure predict launch");
    tk.done();
    return finish_call("closure predict");
}

int msod_closure_sensitivity(int porder, int nelem, int npoin, const int32_t *connec,
                             const double *He, const double *gpvol, const double *dlxigp_ip,
                             const int32_t *invAtoIJK, const int32_t *gmshAtoI,
                             const int32_t *gmshAtoJ, const int32_t *gmshAtoK, double Cp,
                             const double *rho, const double *u, const double *Tem,
                             const double *rerr, float *gnu)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    MsodRange rg_("msod_closure_sensitivity");
    ClosureElemArgs a;
    int st = closure_elem_common(porder, nelem, npoin, connec, He, gpvol, dlxigp_ip, invAtoIJK,
                                 gmshAtoI, gmshAtoJ, gmshAtoK, rho, u, a);
    if (st) return st;
    const int64_t nn = (int64_t)(porder + 1) * (porder + 1) * (porder + 1);
    const void *dT, *de, *dg;
    uint64_t vx;
    if ((st = resolve(Tem, sizeof(double) * npoin, "Tem", &dT, &vx))) return st;
    if ((st = resolve(rerr, sizeof(double) * 4 * (int64_t)npoin, "rerr", &de, &vx))) return st;
    if ((st = resolve(gnu, sizeof(float) * nn * nelem, "gnu", &dg, &vx))) return st;
    a.Tem = (const double *)dT;
    a.Cp = Cp;
    a.rerr = (const double *)de;
    a.gnu = (float *)dg;
    TimedLaunch tk(6);
    HIPCHK(launch_closure_adjoint(porder, a, g.stream), "closure sensitivity launch");
    tk.done();
    return finish_call("closure sensitivity");
}

int msod_closure_gradient(int64_t nsamples, const float *mlp_params, const float *features,
                          const float *gnu, float *grad)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    MsodRange rg_("msod_closure_gradient");
    NEED_INIT();
    if (nsamples <= 0) return fail(MSOD_ERR_ARG, "nsamples=%lld", (long long)nsamples);
    const float *dmlp;
    bool usable;
    int st = closure_weights(mlp_params, &dmlp, &usable);
    if (st) return st;
    if (!usable) return fail(MSOD_ERR_MODEL, "mlp_params missing, not finite or with a negative scale");
    const void *df, *dg, *dgr;
    uint64_t vx;
    if ((st = resolve(features, sizeof(float) * kMlpIn * nsamples, "features", &df, &vx))) return st;
    if ((st = resolve(gnu, sizeof(float) * nsamples, "gnu", &dg, &vx))) return st;
    if ((st = resolve(grad, sizeof(float) * kMlpTrain, "grad", &dgr, &vx))) return st;
    const int nblocks = (int)std::min<int64_t>((nsamples + 63) / 64, 1024);
    void *part;
    if ((st = library_buffer(kScrPartial, sizeof(float) * kMlpTrain * nblocks, &part))) return st;
    TimedLaunch tk(6);
    HIPCHK(launch_mlp_grad(dmlp, (const float *)df, (const float *)dg, nsamples, (float *)part, nblocks,
                           (float *)dgr, g.stream),
           "closure gradient launch");
    tk.done();
    return finish_call("closure gradient");
}

int msod_closure_gradient_noalloc(int64_t nsamples, const float *mlp_params, const float *features,
                                  const float *gnu, float *grad)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    MsodRange rg_("msod_closure_gradient_noalloc");
    NEED_INIT();
    if (nsamples <= 0) return fail(MSOD_ERR_ARG, "nsamples=%lld", (long long)nsamples);
    if (g.device >= kMaxDevices || !g_noalloc_scratch[g.device])
        return fail(MSOD_ERR_STATE, "no partial-sum array on device %d", g.device);
    const int64_t nblocks = std::min<int64_t>((nsamples + 63) / 64, kNoAllocBlocks);
    if (nblocks > kNoAllocBlocks)  // the workgroups stride over the samples: never taken
        return fail(MSOD_ERR_ARG, "closure gradient: %lld workgroups exceed the %d rows of the partial-sum array",
                    (long long)nblocks, kNoAllocBlocks);
    // the weights must already be on the device: a host block would need a
    // library buffer
    const void *dmlp, *df, *dg, *dgr;
    uint64_t vx;
    uintptr_t base;
    if (!mlp_params || !(find_region(mlp_params, &base) || find_region_dev(mlp_params) ||
                         is_device_pointer(mlp_params)))
        return fail(MSOD_ERR_NOT_PRESENT, "mlp_params must be present or device memory");
    int st;
    if ((st = resolve(mlp_params, sizeof(float) * kMlpParams, "mlp_params", &dmlp, &vx))) return st;
    float h[kMlpParams];  // usability check on the host, without a heap buffer
    HIPCHK(hipMemcpyAsync(h, dmlp, sizeof h, hipMemcpyDeviceToHost, g.stream), "reading mlp_params");
    HIPCHK(hipStreamSynchronize(g.stream), "reading mlp_params");
    bool ok = h[kOScale] >= 0.f;
    for (float v : h) ok = ok && std::isfinite(v);
    if (!ok) return fail(MSOD_ERR_MODEL, "mlp_params not finite or with a negative scale");
    if ((st = resolve(features, sizeof(float) * kMlpIn * nsamples, "features", &df, &vx))) return st;
    if ((st = resolve(gnu, sizeof(float) * nsamples, "gnu", &dg, &vx))) return st;
    if ((st = resolve(grad, sizeof(float) * kMlpTrain, "grad", &dgr, &vx))) return st;
    // every workgroup overwrites its whole row before the reduction reads
    // it, so the rows left by an earlier call need no clearing
    HIPCHK(launch_mlp_grad((const float *)dmlp, (const float *)df, (const float *)dg, nsamples,
                           g_noalloc_scratch[g.device], (int)nblocks, (float *)dgr, g.stream),
           "closure gradient launch");
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hipfail(e, "closure gradient");
    if (!g.async) HIPCHK(hipStreamSynchronize(g.stream), "closure gradient");
    return MSOD_OK;
}

int msod_closure_adam(float *mlp_params, const float *grad, float *m, float *v, int step,
                      double lr, double beta1, double beta2, double eps)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    MsodRange rg_("msod_closure_adam");
    NEED_INIT();
    if (step < 1 || !(lr > 0) || !(beta1 >= 0 && beta1 < 1) || !(beta2 >= 0 && beta2 < 1) || !(eps > 0))
        return fail(MSOD_ERR_ARG, "adam: step=%d lr=%g beta1=%g beta2=%g eps=%g", step, lr, beta1, beta2, eps);
    const void *dp, *dg, *dm, *dv;
    uint64_t vx;
    int st;
    if ((st = resolve(mlp_params, sizeof(float) * kMlpParams, "mlp_params", &dp, &vx))) return st;
    if ((st = resolve(grad, sizeof(float) * kMlpTrain, "grad", &dg, &vx))) return st;
    if ((st = resolve(m, sizeof(float) * kMlpTrain, "m", &dm, &vx))) return st;
    if ((st = resolve(v, sizeof(float) * kMlpTrain, "v", &dv, &vx))) return st;
    TimedLaunch tk(6);
    HIPCHK(launch_adam((float *)dp, (const float *)dg, (float *)dm, (float *)dv, step, lr, beta1, beta2,
                       eps, g.stream),
           "adam launch");
    tk.done();
    return finish_call("adam");
}

int msod_closure_loss(int npoin, const double *Rmom, const double *Rener, const double *target,
                      const double *weights, double *rerr, double *loss)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    MsodRange rg_("msod_closure_loss");
    NEED_INIT();
    if (npoin <= 0 || !weights || !loss) return fail(MSOD_ERR_ARG, "npoin=%d, weights and loss needed", npoin);
    const int64_t NP = npoin;
    const void *dmo, *den, *dt, *de;
    uint64_t vx;
    int st;
    if ((st = resolve(Rmom, sizeof(double) * 3 * NP, "Rmom", &dmo, &vx))) return st;
    if ((st = resolve(Rener, sizeof(double) * NP, "Rener", &den, &vx))) return st;
    if ((st = resolve(target, sizeof(double) * 4 * NP, "target", &dt, &vx))) return st;
    if ((st = resolve(rerr, sizeof(double) * 4 * NP, "rerr", &de, &vx))) return st;
    const int nblocks = (int)std::min<int64_t>((NP + 255) / 256, 1024);
    void *buf;
    if ((st = library_buffer(kScrLoss, sizeof(double) * (4 + nblocks), &buf))) return st;
    double *dw = (double *)buf, *lpart = dw + 4;
    std::vector<double> part(nblocks);
    HIPCHK(hipMemcpyAsync(dw, weights, sizeof(double) * 4, hipMemcpyHostToDevice, g.stream), "loss weights");
    TimedLaunch tk(6);
    HIPCHK(launch_residual_error((const double *)dmo, (const double *)den, (const double *)dt, dw, NP,
                                 (double *)de, lpart, nblocks, g.stream),
           "closure loss launch");
    tk.done();
    HIPCHK(hipMemcpyAsync(part.data(), lpart, sizeof(double) * nblocks, hipMemcpyDeviceToHost, g.stream),
           "loss partial sums");
    HIPCHK(hipStreamSynchronize(g.stream), "closure loss");
    double sum = 0.0;
    for (double x : part) sum += x;
    *loss = sum;
    return finish_call("closure loss");
}

}  // extern "C"
