// C ABI of libuclchem_gpu.so (declarations and reference citations: include/uclchem_gpu.h).
// Single translation unit: the kernels are compiled together with their launchers.
#include "ucg_solver.hip"
#include "../../include/uclchem_gpu.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

// Fixed device workspace of ucg_solve_batch_device_noalloc: the work counter (first 256 bytes)
// and the slot workspaces.  Each GPU holds its own copy from the library's first use on it
// until the process ends.  Host code never uses the name itself (on the host it is only a
// stand-in, not a device address): ucg_init records the device address of each GPU's copy.
#define UCG_SCRATCH_BYTES (80u << 20)
static __device__ __align__(256) unsigned char scratch[UCG_SCRATCH_BYTES];

namespace {

std::mutex g_mu;
std::vector<void *> g_scratch;   // per device: the device address of that GPU's scratch
std::string g_err;
bool g_inited = false;
int g_device = -1;
int g_num_cu = 0;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) return fail(UCG_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct NetHandle {
    bool used = false;
    ucg::DevNet dev{};
    std::vector<void *> allocs;
    void *ws = nullptr;
    size_t ws_bytes = 0;
    int *counter = nullptr;
    int *cancel = nullptr;              // device word: highest cancelled launch generation
    int gen = 0;                        // generation of the last launch on this handle
    int cancel_sent = 0;                // highest generation written to *cancel
    ucg::DevNet *d_dev = nullptr;       // device copy of dev, the kernels' table descriptor
    hipStream_t cancel_stream = nullptr;
    int linear_solver = UCG_LINSOL_SPARSE;
};
std::vector<NetHandle> g_nets;

// A request queue: one persistent solve kernel bound to a network that takes requests the
// host posts while it runs.  Requests, the control words and the results live in pinned,
// host-coherent memory the kernel reads and writes directly; the workspace is the queue's
// own, so queues and batch solves on one network never share slots.
struct Queue {
    bool used = false;
    bool drained = false;         // closed and its kernel finished; resources freed when no queue runs
    int handle = -1;
    int capacity = 0, n_slots = 0, lu_bits = 64, next = 0;
    hipStream_t stream = nullptr;
    double *h_params = nullptr;   // [capacity][UCG_NPARAM]
    int *h_ctrl = nullptr;        // [0] requests posted, [1] closed
    int *h_done = nullptr;        // [capacity]
    double *h_y = nullptr;        // [capacity][neq]
    int *h_status = nullptr;      // [capacity]
    int *h_stats = nullptr;       // [capacity][8]
    void *ws = nullptr;
    int *counter = nullptr;
    size_t ws_bytes = 0;
    // shared queue over several networks: their handles, each request's network index
    // (pinned, read by the kernel) and the device array of their descriptors
    std::vector<int> handles;
    int *h_net = nullptr;         // [capacity]
    void *d_nets = nullptr;
    double *h_times = nullptr;    // [2][capacity] device wall clock (ticks) when taken, when finished
    int y_stride = 0;             // doubles between requests' outputs in h_y
};
std::vector<Queue> g_queues;
std::mutex g_cancel_mu;   // orders the cancel-word writes of concurrent ucg_cancel calls

void queue_release(Queue &q)
{
    for (void *p : {(void *)q.h_params, (void *)q.h_ctrl, (void *)q.h_done, (void *)q.h_y, (void *)q.h_status,
                    (void *)q.h_stats, (void *)q.h_net, (void *)q.h_times})
        if (p) (void)hipHostFree(p);
    for (void *p : {q.ws, (void *)q.counter, q.d_nets})
        if (p) (void)hipFree(p);
    if (q.stream) (void)hipStreamDestroy(q.stream);
    q = Queue();
}

// natural cubic spline second derivatives on the host (photoreactions.f90:340-418)
void spline_host(const double *x, const double *y, int n, double *y2)
{
    double u[64];
    y2[0] = 0.0;
    u[0] = 0.0;
    for (int i = 1; i < n - 1; i++) {
        double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
        double p = sig * y2[i - 1] + 2.0;
        y2[i] = (sig - 1.0) / p;
        u[i] = (6.0 * ((y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1])) /
                    (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p;
    }
    y2[n - 1] = 0.0;
    for (int k = n - 2; k >= 0; k--) y2[k] = y2[k] * y2[k + 1] + u[k];
}

int init_locked(int device)
{
    if (g_inited && (device < 0 || device == g_device)) return 0;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(UCG_E_NODEVICE, "no HIP device visible");
    if (device < 0) device = 0;
    if (device >= ndev) return fail(UCG_E_ARG, "device index out of range");
    HIPCHK(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    g_num_cu = prop.multiProcessorCount;
    // photo-rate tables (photoreactions.f90:14-39)
    ucg::PhotoTables t;
    const double lam[30] = {910.0, 950.0, 1000.0, 1050.0, 1110.0, 1180.0, 1250.0, 1390.0, 1490.0, 1600.0,
                            1700.0, 1800.0, 1900.0, 2000.0, 2100.0, 2190.0, 2300.0, 2400.0, 2500.0, 2740.0,
                            3440.0, 4000.0, 4400.0, 5500.0, 7000.0, 9000.0, 12500.0, 22000.0, 34000.0, 1.0e9};
    const double xl[30] = {5.76, 5.18, 4.65, 4.16, 3.73, 3.40, 3.11, 2.74, 2.63, 2.62, 2.54, 2.50, 2.58, 2.78, 3.01,
                           3.12, 2.86, 2.58, 2.35, 2.00, 1.58, 1.42, 1.32, 1.00, 0.75, 0.48, 0.28, 0.12, 0.05, 0.00};
    const double sco[48] = {
        0.000e+00, -1.408e-02, -1.099e-01, -4.400e-01, -1.154e+00, -1.888e+00, -2.760e+00, -4.001e+00,
        -8.539e-02, -1.015e-01, -2.104e-01, -5.608e-01, -1.272e+00, -1.973e+00, -2.818e+00, -4.055e+00,
        -1.451e-01, -1.612e-01, -2.708e-01, -6.273e-01, -1.355e+00, -2.057e+00, -2.902e+00, -4.122e+00,
        -4.559e-01, -4.666e-01, -5.432e-01, -8.665e-01, -1.602e+00, -2.303e+00, -3.146e+00, -4.421e+00,
        -1.303e+00, -1.312e+00, -1.367e+00, -1.676e+00, -2.305e+00, -3.034e+00, -3.758e+00, -5.077e+00,
        -3.883e+00, -3.888e+00, -3.936e+00, -4.197e+00, -4.739e+00, -5.165e+00, -5.441e+00, -6.446e+00};
    for (int i = 0; i < 30; i++) {
        t.lam[i] = lam[i];
        t.xl[i] = xl[i];
    }
    spline_host(t.lam, t.xl, 30, t.xl_d2);
    for (int i = 0; i < 8; i++) t.nco[i] = 12.0 + i;
    for (int i = 0; i < 6; i++) t.nh2[i] = 18.0 + i;
    // the reference hands the 8x6 table to splie2 as 7 rows: its first 42 stored values
    for (int j = 0; j < 7; j++)
        for (int k = 0; k < 6; k++) t.ya[j][k] = sco[j + 7 * k];
    for (int j = 0; j < 7; j++) spline_host(t.nh2, t.ya[j], 6, t.y2a[j]);
    HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(ucg::c_photo), &t, sizeof(t)));
    if ((int)g_scratch.size() < ndev) g_scratch.resize(ndev, nullptr);
    if (!g_scratch[device]) {   // this device's copy of the fixed workspace (device current)
        void *ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail(UCG_E_HIP, "hipGetSymbolAddress(scratch) returned a null address");
        g_scratch[device] = ptr;
    }
    g_device = device;
    g_inited = true;
    return 0;
}

template <typename T>
int upload(NetHandle &h, const T *src, size_t n, const T *&dst)
{
    if (n == 0) {
        dst = nullptr;
        return 0;
    }
    if (!src) return fail(UCG_E_ARG, "null network table");
    void *p = nullptr;
    HIPCHK(hipMalloc(&p, n * sizeof(T)));
    h.allocs.push_back(p);
    HIPCHK(hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice));
    dst = (const T *)p;
    return 0;
}

NetHandle *get(int handle)
{
    if (handle < 0 || handle >= (int)g_nets.size() || !g_nets[handle].used) return nullptr;
    return &g_nets[handle];
}

size_t lds_bytes(const ucg::DevNet &N)
{
    const int npad = (N.neq + 7) & ~7;
    const int region = ucg::flux_region_len(N);
    return sizeof(double) * ((size_t)region + 4 * npad + 16 + ucg::S_NSCAL + ucg::P_NPARAM);
}

int ensure_ws(NetHandle &h, size_t bytes)
{
    if (h.ws_bytes >= bytes) return 0;
    if (h.ws) HIPCHK(hipFree(h.ws));
    h.ws = nullptr;
    h.ws_bytes = 0;
    HIPCHK(hipMalloc(&h.ws, bytes));
    h.ws_bytes = bytes;
    return 0;
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

template <typename T>
int to_dev(DevBuf &b, const T *src, size_t n)
{
    if (!src || n == 0) return 0;
    HIPCHK(hipMalloc(&b.p, n * sizeof(T)));
    HIPCHK(hipMemcpy(b.p, src, n * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}
template <typename T>
int alloc_dev(DevBuf &b, size_t n)
{
    if (n == 0) return 0;
    HIPCHK(hipMalloc(&b.p, n * sizeof(T)));
    return 0;
}

// A caller's buffer for a device-pointer entry point: device (or managed) memory of the
// library's GPU with at least `bytes` bytes between p and the end of its allocation.  Host
// pointers, pinned host memory and other GPUs' memory are refused (the kernels would fault).
int check_dev_buf(const void *p, size_t bytes, const char *name, bool nullable = false)
{
    if (!p) return nullable ? 0 : fail(UCG_E_ARG, std::string("null device buffer ") + name);
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, p) != hipSuccess) {
        (void)hipGetLastError();
        return fail(UCG_E_ARG, std::string(name) + " is not device memory");
    }
    if ((at.type != hipMemoryTypeDevice && at.type != hipMemoryTypeManaged) || at.device != g_device)
        return fail(UCG_E_ARG, std::string(name) + " is not device memory of GPU " + std::to_string(g_device));
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p) == hipSuccess) {
        if ((const char *)p + bytes > (const char *)base + size)
            return fail(UCG_E_ARG, std::string(name) + " is smaller than the batch needs");
    } else {
        (void)hipGetLastError();
    }
    return 0;
}

int check_params(int n, const double *params)
{
    static_assert(UCG_NPARAM == ucg::P_NPARAM, "parameter layouts of the header and the kernels differ");
    for (int p = 0; p < n; p++) {
        const double *q = params + (size_t)p * UCG_NPARAM;
        if (!(q[UCG_P_DENS] > 0) || !(q[UCG_P_TEMP] > 0) || !(q[UCG_P_RELTOL] > 0) || !(q[UCG_P_ABSMIN] >= 0) ||
            !(q[UCG_P_MXSTEP] >= 1) || !(q[UCG_P_ROUT] > q[UCG_P_RIN]) || !(q[UCG_P_MAXSTEPS] >= 0))
            return fail(UCG_E_PARAM, "invalid parameter vector for point " + std::to_string(p));
        int model = (int)q[UCG_P_MODEL];
        if (model != 0 && model != 1 && model != 2)
            return fail(UCG_E_PARAM, "model must be 0 (cloud), 1 (hot core) or 2 (collapse)");
        if (model == 1 && ((int)q[UCG_P_TEMPINDX] < 1 || (int)q[UCG_P_TEMPINDX] > 6))
            return fail(UCG_E_PARAM, "tempindx must be 1..6");
        if (model == 2 && !(q[UCG_P_COLLAPSE] == 1 || q[UCG_P_COLLAPSE] == 2 || q[UCG_P_COLLAPSE] == 3 ||
                            q[UCG_P_COLLAPSE] == 4))   // collapse.f90:43-46 ("unacceptable collapse mode")
            return fail(UCG_E_PARAM, "collapse mode must be 1 (BE1.1), 2 (BE4), 3 (filament) or 4 (ambipolar)");
        if (q[UCG_P_IONMODEL] != 0.0 && q[UCG_P_IONMODEL] != 1.0)
            return fail(UCG_E_PARAM, "ionmodel must be 0 ('L') or 1 ('H')");
        if (q[UCG_P_OMEGA] == 1.0 || q[UCG_P_UVCREFF] == 0.0)
            return fail(UCG_E_PARAM, "omega must differ from 1 and uvcreff from 0");
        // improvedH2CRPDissociation without cosmicRayAttenuation is the reference's physics
        // initialisation error, reported per point (UCG_E_PHYSICS_INIT), not refused here
    }
    return 0;
}

// Bytes of one workgroup slot's workspace: the Newton matrix (FP64 or FP32), the saved
// J-form values of the sparse plan, and the rate coefficients, each 256-byte aligned.
struct SlotLayout {
    size_t mat, jac, stride;
};

SlotLayout slot_layout(const NetHandle &h, int lu_bits)
{
    const ucg::DevNet &N = h.dev;
    const size_t elem = lu_bits == 64 ? sizeof(double) : sizeof(float);
    const bool sparse = h.linear_solver == UCG_LINSOL_SPARSE;
    SlotLayout s;
    s.mat = (((sparse ? (size_t)N.nw_nval : (size_t)N.ld * N.neq) * elem) + 255) & ~(size_t)255;
    s.jac = sparse ? (((size_t)N.nw_nval * sizeof(double) + 255) & ~(size_t)255) : 0;
    s.stride = s.mat + s.jac + (((size_t)N.nreac * sizeof(double) + 255) & ~(size_t)255);
    return s;
}

int launch_kernel(const ucg::DevNet &N, const ucg::DevNet *dN, const ucg::SolveArgs &A, int n_slots, int lu_bits,
                  hipStream_t stream, size_t lds_min = 0)
{
    const size_t lds = std::max(lds_bytes(N), lds_min);
    if (lu_bits == 64) {
        HIPCHK(hipFuncSetAttribute((const void *)ucg::solve_points_kernel<double>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(ucg::solve_points_kernel<double>, dim3(n_slots), dim3(ucg::NT), lds, stream, dN, A);
    } else {
        HIPCHK(hipFuncSetAttribute((const void *)ucg::solve_points_kernel<float>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(ucg::solve_points_kernel<float>, dim3(n_slots), dim3(ucg::NT), lds, stream, dN, A);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

// Optional single-interval inputs and outputs of a solve launch (ucg_interval).
struct IntervalArgs {
    const double *phys_in = nullptr, *t_target = nullptr;
    double *phys_out = nullptr;
};

// A caller-provided workspace for a solve launch: slot workspaces and the work counter.
struct FixedWs {
    void *ws = nullptr;
    int *counter = nullptr;
};

int launch_solve(NetHandle &h, int n, const double *params, const double *y_init, double *y_out, int max_rec,
                 double *rec, int *n_rec, int *status, int *stats, double *t_start, double *t_done, int lu_bits,
                 int n_slots, hipStream_t stream, double *timers = nullptr, const IntervalArgs &iv = IntervalArgs(),
                 const FixedWs &fixed = FixedWs(), double *rates_rec = nullptr)
{
    const ucg::DevNet &N = h.dev;
    if (lu_bits != 64 && lu_bits != 32) return fail(UCG_E_ARG, "lu_bits must be 64 or 32");
    if (n <= 0) return 0;
    if (n_slots <= 0) n_slots = 2 * g_num_cu;
    if (n_slots > n) n_slots = n;
    const bool sparse = h.linear_solver == UCG_LINSOL_SPARSE;
    const SlotLayout sl = slot_layout(h, lu_bits);
    const size_t mat = sl.mat, jac = sl.jac, stride = sl.stride;
    void *ws = fixed.ws;
    int *counter = fixed.counter;
    if (!ws) {
        // sized for the default slot count at least, so later calls with more points do not
        // reallocate between launches
        int rc = ensure_ws(h, stride * (size_t)std::max(n_slots, 2 * g_num_cu));
        if (rc) return rc;
        ws = h.ws;
        counter = h.counter;
    }
    hipLaunchKernelGGL(ucg::solve_init_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, counter, status, n);
    ucg::SolveArgs A = {};
    A.n_points = n;
    A.params = params;
    A.y_init = y_init;
    A.y_out = y_out;
    A.max_rec = rec ? max_rec : 0;
    A.rec = rec;
    A.rates_rec = A.max_rec > 0 ? rates_rec : nullptr;
    A.n_rec = n_rec;
    A.status = status;
    A.stats = stats;
    A.t_start = t_start;
    A.t_done = t_done;
    A.timers = timers;
    A.equilibrate = getenv("UCG_EQUILIBRATE") ? atoi(getenv("UCG_EQUILIBRATE")) : 0;
    A.sparse = sparse ? 1 : 0;
    A.ws = ws;
    A.ws_stride = stride;
    A.ws_mat_bytes = mat;
    A.ws_jac_bytes = jac;
    A.counter = counter;
    A.cancel = h.cancel;
    A.gen = ++h.gen;
    A.phys_in = iv.phys_in;
    A.phys_out = iv.phys_out;
    A.t_target = iv.t_target;
    return launch_kernel(N, h.d_dev, A, n_slots, lu_bits, stream);
}

int piece_common(NetHandle &h, int n, const double *params, const double *phys, ucg::PieceArgs &A,
                 DevBuf &dp, DevBuf &dph)
{
    const ucg::DevNet &N = h.dev;
    if (n <= 0) return fail(UCG_E_ARG, "n must be positive");
    int rc = check_params(n, params);
    if (rc) return rc;
    if ((rc = to_dev(dp, params, (size_t)n * UCG_NPARAM))) return rc;
    if ((rc = to_dev(dph, phys, (size_t)n * 8))) return rc;
    memset(&A, 0, sizeof(A));
    A.n = n;
    A.params = (const double *)dp.p;
    A.phys = (const double *)dph.p;
    const size_t mat = ((size_t)N.ld * N.neq * sizeof(double) + 255) & ~(size_t)255;
    const size_t jac = ((size_t)N.nw_nval * sizeof(double) + 255) & ~(size_t)255;
    const size_t stride = mat + jac + (((size_t)N.nreac * sizeof(double) + 255) & ~(size_t)255);
    if ((rc = ensure_ws(h, stride * n))) return rc;
    A.ws = h.ws;
    A.ws_stride = stride;
    A.ws_mat_bytes = mat;
    A.ws_jac_bytes = jac;
    return 0;
}

}  // namespace

extern "C" {

const char *ucg_last_error(void) { return g_err.c_str(); }

int ucg_init(int device)
{
    std::lock_guard<std::mutex> lk(g_mu);
    return init_locked(device);
}

static int load_tables(const ucg_network_desc *d, NetHandle &h);

// Every index table of the network is range-checked on the host before anything is
// uploaded: species indices in [0, nspec) (or -1 where a slot may be empty), reaction
// indices in [0, nreac), pointer arrays non-decreasing from 0 to their list's length.
static int check_tables(const ucg_network_desc *d)
{
    auto bad = [](const std::string &what) { return fail(UCG_E_ARG, "network table " + what + " out of range"); };
    const long R = d->nreac, Q = d->neq, S = d->nspec;
    if (R <= 0 || S <= 0 || d->nice < 0 || d->nsurfl < 0 || d->nbulkl < 0 || d->nfp < 0 || d->nswap < 0 ||
        d->nloss < 0 || d->nprod < 0 || d->nnz < 0 || d->ncontrib < 0 || d->naff < 0 || d->nsw < 0)
        return fail(UCG_E_ARG, "negative network table size");
    auto in = [](const int *a, long n, long lo, long hi) {   // every a[i] in [lo, hi)
        if (n > 0 && !a) return false;
        for (long i = 0; i < n; i++)
            if (a[i] < lo || a[i] >= hi) return false;
        return true;
    };
    auto mono = [](const int *a, long n, long last) {   // a[0..n] non-decreasing from 0 to last
        if (!a || a[0] != 0 || a[n] != last) return false;
        for (long i = 0; i < n; i++)
            if (a[i + 1] < a[i]) return false;
        return true;
    };
    if (!in(d->rtype, R, 0, ucg::NTYPES)) return bad("rtype");
    if (!in(d->re, 3 * R, -1, S)) return bad("re");
    if (!in(d->pr, 4 * R, -1, S)) return bad("pr");
    if (!in(d->ice_kind, R, 0, 3)) return bad("ice_kind");
    if (!in(d->n_d, R, 0, ucg::MAXFAC + 1)) return bad("n_d");
    if (!in(d->ops, (long)ucg::MAXOPS * R, ucg::OP_DIVSM, S)) return bad("ops");
    if (!d->type_ranges) return bad("type_ranges");
    for (int t = 0; t < ucg::NTYPES; t++) {
        const int a = d->type_ranges[2 * t], b = d->type_ranges[2 * t + 1];
        if (!((a == -1 && b == -1) || (0 <= a && a <= b && b < R))) return bad("type_ranges");
        if (t == ucg::T_FREEZE && a >= 0)
            for (int k = a; k <= b; k++)
                if (d->re[3 * k] < 0) return bad("freeze-out reactant");
    }
    if (!in(d->surface_list, d->nsurfl, 0, S)) return bad("surface_list");
    if (!in(d->bulk_list, d->nbulkl, 0, S)) return bad("bulk_list");
    if (!in(d->bulk_partner_of_surface, Q, -1, S)) return bad("bulk_partner_of_surface");
    if (!in(d->surface_partner_of_bulk, Q, -1, S)) return bad("surface_partner_of_bulk");
    if (!in(d->freeze_partners, d->nfp, 0, R)) return bad("freeze_partners");
    for (long q = 0; q < d->nfp; q++)
        if (d->re[3 * d->freeze_partners[q]] < 0) return bad("freeze-partner reactant");
    if (!in(d->is_surface, Q, 0, 2) || !in(d->is_bulk, Q, 0, 2)) return bad("is_surface/is_bulk");
    if (!in(d->swap_list, d->nswap, 0, R)) return bad("swap_list");
    if (!in(d->jnz_row, d->nnz, 0, Q) || !in(d->jnz_col, d->nnz, 0, Q) || !mono(d->jnz_ptr, d->nnz, d->ncontrib))
        return bad("Jacobian pattern");
    if (!in(d->jc_k, d->ncontrib, 0, R) || !d->jc_m || !d->jc_c) return bad("Jacobian contributions");
    if (!mono(d->aff_ptr, R, d->naff) || !in(d->aff_sp, d->naff, 0, Q) || (d->naff > 0 && !d->aff_cf))
        return bad("affected-species lists");
    if (!mono(d->sw_ptr, Q, d->nsw) || !in(d->sw_k, d->nsw, 0, R) || (d->nsw > 0 && !d->sw_c))
        return bad("totalSwap lists");
    if (!mono(d->loss_ptr, S, d->nloss) || !in(d->loss_idx, d->nloss, 0, R) || !mono(d->prod_ptr, S, d->nprod) ||
        !in(d->prod_idx, d->nprod, 0, R))
        return bad("production/loss lists");
    if (!in(d->nfac, R, 0, 9) || !d->fac) return bad("factor lists");
    for (long k = 0; k < R; k++)
        for (int m = 0; m < d->nfac[k]; m++) {
            const int c = d->fac[ucg::MAXFAC * k + m];
            if (c < ucg::FAC_D || c >= Q) return bad("factor code");
        }
    for (long q = 0; q < d->ncontrib; q++)
        if (d->jc_m[q] < 0 || d->jc_m[q] >= d->nfac[d->jc_k[q]] || d->jc_c[q] < -8 || d->jc_c[q] > 7)
            return bad("Jacobian contribution");
    for (int i = 0; i < 25; i++)
        if (d->special_species[i] < 0 || d->special_species[i] >= Q) return bad("special species index");
    for (int i = 0; i < 10; i++)   // the last, H2 + CRP, may be absent (-1)
        if (d->special_reactions[i] < (i == 9 ? -1 : 0) || d->special_reactions[i] >= R)
            return bad("special reaction index");
    if (d->nion < 0 || !in(d->ion_list, d->nion, 0, S)) return bad("ion_list");
    {   // one parameter row per GAR reaction, paired by position in the GAR range
        const int a = d->type_ranges[2 * ucg::T_GAR], b = d->type_ranges[2 * ucg::T_GAR + 1];
        const long ngar_needed = a < 0 ? 0 : b - a + 1;
        if (d->ngar != ngar_needed || (d->ngar > 0 && !d->gar_params)) return bad("gar_params");
    }
    // the device rate evaluator pairs ranges element-wise (rates.f90:243-279, 104-224)
    const int *tr = d->type_ranges;
    auto len = [&](int t) { return tr[2 * t] < 0 ? 0 : tr[2 * t + 1] - tr[2 * t] + 1; };
    if (len(ucg::T_LH) != len(ucg::T_LHDES)) return fail(UCG_E_ARG, "LH and LHDES ranges differ in length");
    if (len(ucg::T_ER) != len(ucg::T_ERDES)) return fail(UCG_E_ARG, "ER and ERDES ranges differ in length");
    for (int t : {ucg::T_DESOH2, ucg::T_DESCR, ucg::T_DEUVCR, ucg::T_THERM})
        if (len(t) > 1 && len(t) < d->nfp) return fail(UCG_E_ARG, "desorption range shorter than freeze partners");
    return 0;
}

// Every index of the sparse Newton plan is range-checked here, on the host, so that no
// table handed in by a caller can make the device address outside its buffers.
static int check_plan(const ucg_network_desc *d)
{
    const int *z = d->nw_sizes;
    const long NN = z[0], ns = z[1], T = z[2], nsp = z[3], nlev = z[4], ntg = z[5], npair = z[6], nph = z[7];
    const long nphs = z[8], ngs = z[9];
    const long nval = nsp + T * T;
    auto bad = [](const char *what) { return fail(UCG_E_ARG, std::string("sparse Newton plan: ") + what); };
    if (NN != d->neq + 2 || ns < 0 || ns > NN || T != NN - ns || nsp < 0 || nlev < 0 || ntg < 0 || npair < 0 ||
        nph < 1 || nphs < 0 || ngs < 0)
        return bad("inconsistent sizes");
    if (T > 0 && NN - ns > ucg::NT * ucg::EPT) return bad("tail too large");
    // the tail's Gauss-Jordan keeps its 8x8 pivot-block inverse in the right-hand-side vector
    if (T > 0 && ((d->neq + 7) & ~7) < ucg::GJ_B * ucg::GJ_B) return bad("network too small for the dense tail");
    auto in = [](const int *a, long n, long lo, long hi) {   // every a[i] in [lo, hi)
        if (n > 0 && !a) return false;
        for (long i = 0; i < n; i++)
            if (a[i] < lo || a[i] >= hi) return false;
        return true;
    };
    auto mono = [](const int *a, long n, long first, long last) {   // a[0..n], non-decreasing
        if (!a || a[0] != first || a[n] != last) return false;
        for (long i = 0; i < n; i++)
            if (a[i + 1] < a[i]) return false;
        return true;
    };
    if (!in(d->nw_perm, NN, 0, NN) || !in(d->nw_inv, NN, 0, NN)) return bad("permutation out of range");
    for (long i = 0; i < NN; i++)
        if (d->nw_perm[d->nw_inv[i]] != i) return bad("perm and inv are not inverse");
    if (!mono(d->nw_sp_ptr, NN, 0, nsp) || !in(d->nw_sp_col, nsp, 0, NN)) return bad("sparse rows");
    for (long r = 0; r < ns; r++)
        if (d->nw_sp_dpos[r] < d->nw_sp_ptr[r] || d->nw_sp_dpos[r] >= d->nw_sp_ptr[r + 1] ||
            d->nw_sp_col[d->nw_sp_dpos[r]] != r)
            return bad("diagonal positions");
    if (!mono(d->nw_lev_ptr, nlev, 0, ns) || !in(d->nw_lev_rows, ns, 0, ns)) return bad("levels");
    if (!mono(d->nw_ph_ptr, nph, 0, ntg) || !in(d->nw_tg_slot, ntg, 0, nval) || !mono(d->nw_tg_ptr, ntg, 0, npair) ||
        !in(d->nw_tg_div, ntg, -1, nval) || !in(d->nw_pr_l, npair, 0, nval) || !in(d->nw_pr_u, npair, 0, nval))
        return bad("elimination schedule");
    const long Q = d->neq;
    if (!in(d->nw_jnz_slot, d->nnz, 0, nval) || !in(d->nw_diag_slot, Q, -1, nval) ||
        !in(d->nw_ph_slot, nphs, 0, nval) || !in(d->nw_tr_slot, 8L * d->nsurfl, 0, nval) ||
        !in(d->nw_s1c, Q, -1, nval) || !in(d->nw_s1r, Q, -1, nval) || !in(d->nw_s2c, Q, -1, nval) ||
        !in(d->nw_s2r, Q, -1, nval) || !mono(d->nw_gs_ptr, Q, 0, ngs) || !in(d->nw_gs_slot, ngs, 0, nval) ||
        !in(d->nw_misc, 5, 0, nval) || !in(d->nw_nbrow, d->nbulkl, 0, nval) || !in(d->nw_nsrow, d->nsurfl, 0, nval))
        return bad("assembly slots");
    // charge row: present whenever the network has positive ions; its ion slots lie in it
    if (d->nw_nerow < 0 || (d->nion > 0 && d->nw_nerow == 0) || !in(d->nw_erow, d->nw_nerow, 0, nval) ||
        !in(d->nw_eion, d->nion, 0, nval))
        return bad("charge row slots");
    for (int q = 0; q < d->nion; q++) {
        bool found = false;
        for (int r = 0; r < d->nw_nerow && !found; r++) found = d->nw_erow[r] == d->nw_eion[q];
        if (!found) return bad("charge row slots");
    }
    long naff_photo = 0;
    for (int k : {d->special_reactions[6], d->special_reactions[7]}) {
        if (k < 0 || k >= d->nreac) return bad("photo reactions");
        naff_photo += d->aff_ptr[k + 1] - d->aff_ptr[k];
    }
    if (nphs != 3 * naff_photo) return bad("photo slots");
    return 0;
}

int ucg_network_load(const ucg_network_desc *d, int *handle)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!d || !handle) return fail(UCG_E_ARG, "null argument");
    if (d->neq != d->nspec + 1 || d->neq > ucg::NT * ucg::EPT)
        return fail(UCG_E_ARG, "network too large: NEQ must be <= " + std::to_string(ucg::NT * ucg::EPT));
    // tables are validated on the host before the device is touched
    int rc = check_tables(d);
    if (rc || (rc = check_plan(d))) return rc;
    if ((rc = init_locked(-1))) return rc;
    NetHandle h;
    h.used = true;
    if ((rc = load_tables(d, h))) {   // release whatever was uploaded before the failure
        for (void *p : h.allocs) (void)hipFree(p);
        return rc;
    }
    g_nets.push_back(h);
    *handle = (int)g_nets.size() - 1;
    return 0;
}

// Reads the flat descriptor file written by makerates_hip.save_network_bin.
int ucg_network_load_file(const char *path, int *handle)
{
    if (!path || !handle) return fail(UCG_E_ARG, "null argument");
    FILE *f = fopen(path, "rb");
    if (!f) return fail(UCG_E_ARG, std::string("cannot open network file ") + path);
    std::vector<char> buf;
    {
        char tmp[1 << 16];
        size_t r;
        while ((r = fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + r);
        fclose(f);
    }
    if (buf.size() < 12 || memcmp(buf.data(), "UCGNET01", 8) != 0)
        return fail(UCG_E_ARG, std::string("not a UCGNET01 network file: ") + path);
    int32_t nrec;
    memcpy(&nrec, buf.data() + 8, 4);
    size_t off = 12;
    struct Rec { int dtype; int64_t count; const char *data; };
    std::vector<std::pair<std::string, Rec>> recs;
    for (int r = 0; r < nrec; r++) {
        if (off + 36 > buf.size()) return fail(UCG_E_ARG, "truncated network file");
        char name[25] = {0};
        memcpy(name, buf.data() + off, 24);
        Rec rc;
        memcpy(&rc.dtype, buf.data() + off + 24, 4);
        memcpy(&rc.count, buf.data() + off + 28, 8);
        off += 36;
        const size_t esz = rc.dtype == 0 ? 8 : rc.dtype == 1 ? 4 : rc.dtype == 2 ? 1 : 0;
        if (!esz || rc.count < 0 || off + esz * (size_t)rc.count > buf.size())
            return fail(UCG_E_ARG, std::string("bad record ") + name + " in network file");
        rc.data = buf.data() + off;
        off += esz * (size_t)rc.count;
        off += (8 - (esz * (size_t)rc.count) % 8) % 8;
        recs.push_back({name, rc});
    }
    auto find = [&](const char *name, int dtype, int64_t need, const void **out) -> int {
        for (auto &kv : recs)
            if (kv.first == name) {
                if (kv.second.dtype != dtype || (need >= 0 && kv.second.count != need))
                    return fail(UCG_E_ARG, std::string("network file record ") + name + " has the wrong type or size");
                *out = kv.second.data;
                return 0;
            }
        return fail(UCG_E_ARG, std::string("network file lacks record ") + name);
    };
    ucg_network_desc d;
    memset(&d, 0, sizeof d);
    const void *p;
    int rc;
    if ((rc = find("sizes", 1, 14, &p))) return rc;
    int32_t sz[14];
    memcpy(sz, p, sizeof sz);
    d.nspec = sz[0]; d.neq = sz[1]; d.nreac = sz[2]; d.nice = sz[3]; d.nsurfl = sz[4]; d.nbulkl = sz[5];
    d.nfp = sz[6]; d.nswap = sz[7]; d.nloss = sz[8]; d.nprod = sz[9]; d.nnz = sz[10]; d.ncontrib = sz[11];
    d.naff = sz[12]; d.nsw = sz[13];
    // copies keep the int32/double arrays aligned whatever the file offsets are
    std::vector<std::vector<double>> keep_d;
    std::vector<std::vector<int>> keep_i;
    std::vector<std::vector<signed char>> keep_c;
    auto getd = [&](const char *nm, int64_t n, const double **dst) -> int {
        int r = find(nm, 0, n, &p);
        if (r) return r;
        keep_d.emplace_back((size_t)n);
        memcpy(keep_d.back().data(), p, 8 * (size_t)n);
        *dst = keep_d.back().data();
        return 0;
    };
    auto geti = [&](const char *nm, int64_t n, const int **dst) -> int {
        int r = find(nm, 1, n, &p);
        if (r) return r;
        keep_i.emplace_back((size_t)n);
        memcpy(keep_i.back().data(), p, 4 * (size_t)n);
        *dst = keep_i.back().data();
        return 0;
    };
    const int64_t R = d.nreac, Q = d.neq, S = d.nspec;
    keep_d.reserve(8);
    keep_i.reserve(40);
#define GD(f, n) if ((rc = getd(#f, (n), &d.f))) return rc
#define GI(f, n) if ((rc = geti(#f, (n), &d.f))) return rc
    GD(alpha, R); GD(beta, R); GD(gama, R); GD(tmin, R); GD(tmax, R); GD(mass, S); GD(rstat, 5 * R);
    GI(rtype, R); GI(re, 3 * R); GI(pr, 4 * R); GI(ice_kind, R); GI(n_d, R); GI(ops, ucg::MAXOPS * R);
    GI(fac, ucg::MAXFAC * R); GI(nfac, R); GI(type_ranges, 2 * ucg::NTYPES);
    GI(loss_ptr, S + 1); GI(loss_idx, d.nloss); GI(prod_ptr, S + 1); GI(prod_idx, d.nprod);
    GI(swap_list, d.nswap); GI(surface_list, d.nsurfl); GI(bulk_list, d.nbulkl);
    GI(bulk_partner_of_surface, Q); GI(surface_partner_of_bulk, Q); GI(freeze_partners, d.nfp);
    GI(is_surface, Q); GI(is_bulk, Q); GI(jnz_row, d.nnz); GI(jnz_col, d.nnz); GI(jnz_ptr, d.nnz + 1);
    GI(jc_k, d.ncontrib); GI(jc_m, d.ncontrib); GI(jc_c, d.ncontrib); GI(aff_ptr, R + 1);
    GI(aff_sp, d.naff); GI(aff_cf, d.naff); GI(sw_ptr, Q + 1); GI(sw_k, d.nsw); GI(sw_c, d.nsw);
#undef GD
#undef GI
    if ((rc = find("extrapolate", 2, R, &p))) return rc;
    keep_c.emplace_back((const signed char *)p, (const signed char *)p + R);
    d.extrapolate = keep_c.back().data();
    if ((rc = find("special_species", 1, 25, &p))) return rc;
    memcpy(d.special_species, p, sizeof d.special_species);
    if ((rc = find("special_reactions", 1, 10, &p))) return rc;
    memcpy(d.special_reactions, p, sizeof d.special_reactions);
    if ((rc = find("nion", 1, 1, &p))) return rc;
    memcpy(&d.nion, p, sizeof d.nion);
    if (d.nion < 0) return fail(UCG_E_ARG, "network file record nion is negative");
    if ((rc = geti("ion_list", d.nion, &d.ion_list))) return rc;
    if ((rc = find("ngar", 1, 1, &p))) return rc;
    memcpy(&d.ngar, p, sizeof d.ngar);
    if (d.ngar < 0) return fail(UCG_E_ARG, "network file record ngar is negative");
    if ((rc = getd("gar_params", 7 * (int64_t)d.ngar, &d.gar_params))) return rc;
    if ((rc = find("nw_sizes", 1, 10, &p))) return rc;
    memcpy(d.nw_sizes, p, sizeof d.nw_sizes);
    {
        const int *z = d.nw_sizes;
        const int64_t NN = z[0], ns = z[1], nsp = z[3], nlev = z[4], ntg = z[5], npair = z[6], nph = z[7];
        keep_i.reserve(keep_i.size() + 30);
#define GI(f, n) if ((rc = geti(#f, (n), &d.f))) return rc
        GI(nw_perm, NN); GI(nw_inv, NN); GI(nw_sp_ptr, NN + 1); GI(nw_sp_col, nsp); GI(nw_sp_dpos, ns);
        GI(nw_lev_ptr, nlev + 1); GI(nw_lev_rows, ns); GI(nw_ph_ptr, nph + 1); GI(nw_tg_slot, ntg);
        GI(nw_tg_ptr, ntg + 1); GI(nw_tg_div, ntg); GI(nw_pr_l, npair); GI(nw_pr_u, npair);
        GI(nw_jnz_slot, d.nnz); GI(nw_diag_slot, Q); GI(nw_ph_slot, z[8]); GI(nw_tr_slot, 8 * (int64_t)d.nsurfl);
        GI(nw_s1c, Q); GI(nw_s1r, Q); GI(nw_s2c, Q); GI(nw_s2r, Q); GI(nw_gs_ptr, Q + 1); GI(nw_gs_slot, z[9]);
        GI(nw_misc, 5); GI(nw_nbrow, d.nbulkl); GI(nw_nsrow, d.nsurfl);
        if ((rc = find("nw_nerow", 1, 1, &p))) return rc;
        memcpy(&d.nw_nerow, p, sizeof d.nw_nerow);
        if (d.nw_nerow < 0) return fail(UCG_E_ARG, "network file record nw_nerow is negative");
        GI(nw_erow, d.nw_nerow); GI(nw_eion, d.nion);
#undef GI
    }
    return ucg_network_load(&d, handle);
}

// Uploads a network whose tables check_tables and check_plan have accepted.
static int load_tables(const ucg_network_desc *d, NetHandle &h)
{
    int rc;
    ucg::DevNet &N = h.dev;
    N.nspec = d->nspec; N.neq = d->neq; N.nreac = d->nreac; N.nice = d->nice; N.nsurfl = d->nsurfl;
    N.nbulkl = d->nbulkl; N.nfp = d->nfp; N.nswap = d->nswap; N.nnz = d->nnz;
    N.ld = (d->neq + 15) & ~15;
    const size_t R = d->nreac, Q = d->neq;
#define UP(field, count) if ((rc = upload(h, d->field, (count), N.field))) return rc
    UP(alpha, R); UP(beta, R); UP(gama, R); UP(tmin, R); UP(tmax, R); UP(mass, (size_t)d->nspec);
    UP(rstat, 5 * R); UP(rtype, R); UP(re, 3 * R); UP(pr, 4 * R); UP(ice_kind, R); UP(n_d, R);
    UP(ops, ucg::MAXOPS * R); UP(fac, ucg::MAXFAC * R); UP(nfac, R); UP(extrapolate, R);
    {   // factor records for the term partials: int16 codes x 8, offsets, packed contributions
        std::vector<int> facp(4 * R, 0), foff(R + 1, 0), jpk((size_t)d->ncontrib);
        for (size_t k = 0; k < R; k++) {
            const int nf = d->nfac[k];
            if (nf < 0 || nf > 8) return fail(UCG_E_ARG, "a reaction term has more than 8 factors");
            foff[k + 1] = foff[k] + nf;
            for (int m = 0; m < 8; m++) {
                const int c = m < nf ? d->fac[ucg::MAXFAC * k + m] : ucg::OP_NONE;
                if (c < ucg::FAC_D || c >= d->neq) return fail(UCG_E_ARG, "factor code out of range");
                facp[4 * k + m / 2] |= (int)((unsigned)(c & 0xffff) << (16 * (m & 1)));
            }
        }
        for (int q = 0; q < d->ncontrib; q++) {
            const int k = d->jc_k[q], m = d->jc_m[q], c = d->jc_c[q];
            if (k < 0 || k >= d->nreac || m < 0 || m >= d->nfac[k] || c < -8 || c > 7)
                return fail(UCG_E_ARG, "Jacobian contribution out of range");
            jpk[q] = ((foff[k] + m) << 4) | (c + 8);
        }
        N.npart = foff[R];
        N.npart_pad = (N.npart + 7) & ~7;
        if ((rc = upload(h, facp.data(), facp.size(), N.facp))) return rc;
        if ((rc = upload(h, foff.data(), foff.size(), N.fac_off))) return rc;
        if ((rc = upload(h, jpk.data(), jpk.size(), N.jc_pk))) return rc;
    }
    {   // Jacobian entries in two balanced lists: short contribution lists one per lane, longest
        // first and dealt to the lanes in alternating directions (so each lane's total is near
        // the mean), long ones (the density column of the hydrogen-bearing species: hundreds of
        // terms) by groups of JZ_GROUP lanes
        std::vector<int> sh, lo;
        for (int n = 0; n < d->nnz; n++)
            (d->jnz_ptr[n + 1] - d->jnz_ptr[n] > ucg::JZ_LONG ? lo : sh).push_back(n);
        std::stable_sort(sh.begin(), sh.end(), [&](int a, int b) {
            return d->jnz_ptr[a + 1] - d->jnz_ptr[a] > d->jnz_ptr[b + 1] - d->jnz_ptr[b];
        });
        std::vector<int> dealt(sh.size());
        const size_t nt = ucg::NT, full = sh.size() / nt * nt;
        for (size_t p = 0; p < sh.size(); p++) {
            const size_t blk = p / nt, r = p % nt;
            const size_t lane = (p < full && (blk & 1)) ? nt - 1 - r : r;
            dealt[blk * nt + lane] = sh[p];
        }
        N.nj_short = (int)dealt.size();
        N.nj_long = (int)lo.size();
        if ((rc = upload(h, dealt.data(), dealt.size(), N.jz_short))) return rc;
        if ((rc = upload(h, lo.data(), lo.size(), N.jz_long))) return rc;
    }
    {   // production/loss lists index fluxes: checked here, the device trusts them
        for (int kind = 0; kind < 2; kind++) {
            const int *ptr = kind ? d->loss_ptr : d->prod_ptr, *idx = kind ? d->loss_idx : d->prod_idx;
            const int len = kind ? d->nloss : d->nprod;
            if (ptr[0] != 0 || ptr[d->nspec] != len) return fail(UCG_E_ARG, "production/loss list bounds");
            for (int i = 0; i < d->nspec; i++)
                if (ptr[i + 1] < ptr[i]) return fail(UCG_E_ARG, "production/loss lists not ordered");
            for (int q = 0; q < len; q++)
                if (idx[q] < 0 || idx[q] >= d->nreac) return fail(UCG_E_ARG, "production/loss reaction index");
        }
    }
    {   // balanced production/loss sum tasks (chunks of at most GT_CHUNK flux indices)
        const int GT_CHUNK = 24;
        struct Task { int start, end, species, kind; };
        std::vector<Task> tasks;
        for (int i = 0; i < d->nspec; i++)
            for (int kind = 0; kind < 2; kind++) {
                const int *ptr = kind ? d->loss_ptr : d->prod_ptr;
                for (int a = ptr[i]; a < ptr[i + 1]; a += GT_CHUNK)
                    tasks.push_back({a, std::min(a + GT_CHUNK, ptr[i + 1]), i, kind});
            }
        std::vector<int> order(tasks.size());
        for (size_t t = 0; t < order.size(); t++) order[t] = (int)t;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
            return tasks[x].end - tasks[x].start > tasks[y].end - tasks[y].start;
        });
        std::vector<int> rec(4 * tasks.size()), where(tasks.size());
        for (size_t t = 0; t < order.size(); t++) {
            const Task &k = tasks[order[t]];
            rec[4 * t] = k.start; rec[4 * t + 1] = k.end; rec[4 * t + 2] = k.species; rec[4 * t + 3] = k.kind;
            where[order[t]] = (int)t;
        }
        std::vector<int> sptr(2 * (size_t)d->nspec + 1, 0), stask;
        for (size_t t = 0; t < tasks.size(); t++) sptr[2 * tasks[t].species + tasks[t].kind + 1]++;
        for (size_t q = 1; q < sptr.size(); q++) sptr[q] += sptr[q - 1];
        stask.resize(tasks.size());
        std::vector<int> fill(sptr.begin(), sptr.end() - 1);
        for (size_t t = 0; t < tasks.size(); t++)   // tasks were generated in chunk order
            stask[fill[2 * tasks[t].species + tasks[t].kind]++] = where[t];
        N.ngt = (int)tasks.size();
        N.gt_off = (d->nreac + 7) & ~7;
        if ((rc = upload(h, rec.data(), rec.size(), N.gt_rec))) return rc;
        if ((rc = upload(h, sptr.data(), sptr.size(), N.gt_sp_ptr))) return rc;
        if ((rc = upload(h, stask.data(), stask.size(), N.gt_sp_task))) return rc;
    }
    {   // packed term records {n_d, ops[6], nfac}
        std::vector<int> prog(8 * R);
        for (size_t k = 0; k < R; k++) {
            prog[8 * k] = d->n_d[k];
            for (int q = 0; q < ucg::MAXOPS; q++) prog[8 * k + 1 + q] = d->ops[ucg::MAXOPS * k + q];
            prog[8 * k + 7] = d->nfac[k];
        }
        if ((rc = upload(h, prog.data(), prog.size(), N.prog))) return rc;
    }
    UP(type_ranges, (size_t)2 * ucg::NTYPES);
    UP(loss_ptr, (size_t)d->nspec + 1); UP(loss_idx, (size_t)d->nloss);
    UP(prod_ptr, (size_t)d->nspec + 1); UP(prod_idx, (size_t)d->nprod); UP(swap_list, (size_t)d->nswap);
    UP(surface_list, (size_t)d->nsurfl); UP(bulk_list, (size_t)d->nbulkl);
    UP(bulk_partner_of_surface, Q); UP(surface_partner_of_bulk, Q);
    UP(freeze_partners, (size_t)d->nfp); UP(is_surface, Q); UP(is_bulk, Q);
    UP(jnz_row, (size_t)d->nnz); UP(jnz_col, (size_t)d->nnz); UP(jnz_ptr, (size_t)d->nnz + 1);
    UP(jc_k, (size_t)d->ncontrib); UP(jc_m, (size_t)d->ncontrib); UP(jc_c, (size_t)d->ncontrib);
    UP(aff_ptr, R + 1); UP(aff_sp, (size_t)d->naff); UP(aff_cf, (size_t)d->naff);
    UP(sw_ptr, Q + 1); UP(sw_k, (size_t)d->nsw); UP(sw_c, (size_t)d->nsw);
    {   // the Jacobian sums d totalSwap/dy per bulk species without ordering: one BULKSWAP each
        std::vector<char> seen((size_t)d->neq, 0);
        for (int q = 0; q < d->nswap; q++) {
            const int k = d->swap_list[q];
            if (k < 0 || k >= d->nreac) return fail(UCG_E_ARG, "swap list index out of range");
            const int b = d->ops[ucg::MAXOPS * k];
            if (b < 0 || b >= d->neq || seen[b]) return fail(UCG_E_ARG, "BULKSWAP reactants must be distinct species");
            seen[b] = 1;
        }
    }
    {
        const int *z = d->nw_sizes;
        N.nw_N = z[0]; N.nw_ns = z[1]; N.nw_T = z[2]; N.nw_nsp = z[3]; N.nw_nlev = z[4]; N.nw_nph = z[7];
        N.nw_nval = z[3] + z[2] * z[2];
        N.nw_stage = 1;   // cleared below when the staged layout exceeds the device's LDS
        const size_t NN = z[0], ns = z[1], nsp = z[3], nlev = z[4], ntg = z[5], npair = z[6], nph = z[7];
        UP(nw_perm, NN); UP(nw_inv, NN); UP(nw_sp_ptr, NN + 1); UP(nw_sp_col, nsp); UP(nw_sp_dpos, ns);
        UP(nw_lev_ptr, nlev + 1); UP(nw_lev_rows, ns); UP(nw_ph_ptr, nph + 1);
        // targets of each elimination phase, longest pair list first: a pass of the workgroup
        // over a phase then takes the long chains together instead of one per pass
        std::vector<int> tslot(ntg), tptr(ntg + 1, 0), tdiv(ntg), pl(npair), pu(npair);
        for (size_t p = 0; p < nph; p++) {
            std::vector<int> ord;
            for (int t = d->nw_ph_ptr[p]; t < d->nw_ph_ptr[p + 1]; t++) ord.push_back(t);
#ifndef UCG_SORT_TARGETS
#define UCG_SORT_TARGETS 1
#endif
            if (UCG_SORT_TARGETS)
                std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
                    return d->nw_tg_ptr[x + 1] - d->nw_tg_ptr[x] > d->nw_tg_ptr[y + 1] - d->nw_tg_ptr[y];
                });
            int t2 = d->nw_ph_ptr[p];
            for (int t : ord) {
                tslot[t2] = d->nw_tg_slot[t];
                tdiv[t2] = d->nw_tg_div[t];
                int q2 = tptr[t2];
                for (int q = d->nw_tg_ptr[t]; q < d->nw_tg_ptr[t + 1]; q++, q2++) {
                    pl[q2] = d->nw_pr_l[q];
                    pu[q2] = d->nw_pr_u[q];
                }
                tptr[t2 + 1] = q2;
                t2++;
            }
        }
        if ((rc = upload(h, tslot.data(), tslot.size(), N.nw_tg_slot))) return rc;
        if ((rc = upload(h, tptr.data(), tptr.size(), N.nw_tg_ptr))) return rc;
        if ((rc = upload(h, tdiv.data(), tdiv.size(), N.nw_tg_div))) return rc;
        UP(nw_jnz_slot, (size_t)d->nnz); UP(nw_diag_slot, Q); UP(nw_ph_slot, (size_t)z[8]);
        UP(nw_tr_slot, (size_t)8 * d->nsurfl); UP(nw_s1c, Q); UP(nw_s1r, Q); UP(nw_s2c, Q); UP(nw_s2r, Q);
        UP(nw_gs_ptr, Q + 1); UP(nw_gs_slot, (size_t)z[9]); UP(nw_misc, (size_t)5);
        UP(nw_nbrow, (size_t)d->nbulkl); UP(nw_nsrow, (size_t)d->nsurfl);
        N.nw_nerow = d->nw_nerow;
        UP(nw_erow, (size_t)d->nw_nerow); UP(nw_eion, (size_t)d->nion);
        // solve row records: sparse rows in level order, then the tail rows
        std::vector<int> sv(4 * NN);
        for (size_t q = 0; q < NN; q++) {
            const int r = q < ns ? d->nw_lev_rows[q] : (int)q;
            sv[4 * q] = r;
            sv[4 * q + 1] = d->nw_sp_ptr[r];
            sv[4 * q + 2] = r < (int)ns ? d->nw_sp_dpos[r] : d->nw_sp_ptr[r + 1];
            sv[4 * q + 3] = d->nw_sp_ptr[r + 1];
        }
        if ((rc = upload(h, sv.data(), sv.size(), N.sv_rec))) return rc;
        std::vector<int> pr(2 * npair);
        for (size_t q = 0; q < npair; q++) {
            pr[2 * q] = pl[q];
            pr[2 * q + 1] = pu[q];
        }
        if ((rc = upload(h, pr.data(), pr.size(), N.nw_pr))) return rc;
        {   // sources of the last phase when it is the pure Schur update of the tail
            const int *pp = d->nw_ph_ptr;
            const long t0 = pp[nph - 1], t1 = pp[nph];
            const long q0 = tptr[t0], q1 = tptr[t1];
            bool ok = nph >= 2 && q1 > q0;
            for (long t = t0; t < t1 && ok; t++) ok = tslot[t] >= (int)nsp && tdiv[t] < 0;
            for (long q = q0; q < q1 && ok; q++) ok = pl[q] < (int)nsp && pu[q] < (int)nsp;
            N.nlp_src = 0;
            if (ok) {
                std::vector<int> local(nsp, -1), src, lpp(2 * (q1 - q0));
                for (long q = q0; q < q1; q++)
                    for (int h2 = 0; h2 < 2; h2++) {
                        const int sl = h2 ? pu[q] : pl[q];
                        if (local[sl] < 0) {
                            local[sl] = (int)src.size();
                            src.push_back(sl);
                        }
                        lpp[2 * (q - q0) + h2] = local[sl];
                    }
                N.nlp_src = (int)src.size();
                if ((rc = upload(h, src.data(), src.size(), N.lp_src))) return rc;
                if ((rc = upload(h, lpp.data(), lpp.size(), N.lp_pr))) return rc;
            }
        }
        std::vector<signed char> uns((size_t)N.nw_nval, 0);
        for (size_t i = 0; i < Q; i++) {
            if (d->nw_s1c[i] >= 0) uns[d->nw_s1c[i]] = 1;
            if (d->nw_s2c[i] >= 0) uns[d->nw_s2c[i]] = 1;
        }
        uns[d->nw_misc[2]] = 1;
        const int8_t *du = nullptr;
        if ((rc = upload(h, (const int8_t *)uns.data(), uns.size(), du))) return rc;
        N.nw_unscaled = du;
    }
#undef UP
    const int *ss = d->special_species;
    N.nh = ss[0]; N.nh2 = ss[1]; N.nhe = ss[2]; N.nc = ss[3]; N.ncx = ss[4]; N.nn = ss[5]; N.no = ss[6];
    N.nf = ss[7]; N.np = ss[8]; N.nmg = ss[9]; N.nmgx = ss[10]; N.nna = ss[11]; N.nli = ss[12];
    N.npah = ss[13]; N.nsx = ss[14]; N.nsix = ss[15]; N.nclx = ss[16]; N.nd = ss[17]; N.n18o = ss[18];
    N.n15n = ss[19]; N.n13c = ss[20]; N.nelec = ss[21]; N.nco = ss[22]; N.nsurface = ss[23]; N.nbulk = ss[24];
    const int *sr = d->special_reactions;
    N.r_h2form_ct = sr[0]; N.r_h2form_er = sr[1]; N.r_h2form_erdes = sr[2]; N.r_hfreeze = sr[3];
    N.r_efreeze = sr[4]; N.r_h2freeze = sr[5]; N.r_h2_hv = sr[6]; N.r_co_hv = sr[7]; N.r_c_hv = sr[8];
    N.r_h2_crp = sr[9];
    N.nion = d->nion;
    if ((rc = upload(h, d->ion_list, (size_t)d->nion, N.ion_list))) return rc;
    N.ngar = d->ngar;
    if ((rc = upload(h, d->gar_params, 7 * (size_t)d->ngar, N.gar))) return rc;
    HIPCHK(hipMalloc(&h.counter, sizeof(int)));
    h.allocs.push_back(h.counter);
    HIPCHK(hipMalloc(&h.cancel, sizeof(int)));
    h.allocs.push_back(h.cancel);
    HIPCHK(hipMemset(h.cancel, 0, sizeof(int)));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, g_device));
    if (lds_bytes(N) > (size_t)prop.sharedMemPerBlock) N.nw_stage = 0;   // factors stay in the workspace
    size_t lds = lds_bytes(N);
    if (lds > (size_t)prop.sharedMemPerBlock)
        return fail(UCG_E_ARG, "network needs " + std::to_string(lds) + " B of LDS per workgroup, device allows " +
                                   std::to_string(prop.sharedMemPerBlock));
    HIPCHK(hipMalloc(&h.d_dev, sizeof(ucg::DevNet)));   // the descriptor the kernels read (final values)
    h.allocs.push_back(h.d_dev);
    HIPCHK(hipMemcpy(h.d_dev, &N, sizeof(ucg::DevNet), hipMemcpyHostToDevice));
    HIPCHK(hipStreamCreateWithFlags(&h.cancel_stream, hipStreamNonBlocking));   // last: nothing fails after it
    return 0;
}

int ucg_network_free(int handle)
{
    std::lock_guard<std::mutex> lk(g_mu);
    NetHandle *h = get(handle);
    if (!h) return fail(UCG_E_HANDLE, "bad network handle");
    for (const Queue &q : g_queues)
        if (q.used && !q.drained && std::find(q.handles.begin(), q.handles.end(), handle) != q.handles.end())
            return fail(UCG_E_ARG, "close the network's request queues first");
    for (void *p : h->allocs) (void)hipFree(p);
    if (h->ws) (void)hipFree(h->ws);
    if (h->cancel_stream) (void)hipStreamDestroy(h->cancel_stream);
    *h = NetHandle();
    return 0;
}

int ucg_cancel(int handle)
{
    // g_cancel_mu orders concurrent cancels; g_mu is held only to read the handle, so that a
    // cancel does not wait behind a host-buffer solve that holds it for its whole run
    std::lock_guard<std::mutex> ck(g_cancel_mu);
    int *word;
    hipStream_t stream;
    int gen;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        NetHandle *h = get(handle);
        if (!h) return fail(UCG_E_HANDLE, "bad network handle");
        if (h->gen <= h->cancel_sent) return 0;   // nothing launched since the last cancel
        word = h->cancel;
        stream = h->cancel_stream;
        gen = h->gen;
        h->cancel_sent = gen;
    }
    HIPCHK(hipMemcpyAsync(word, &gen, sizeof(int), hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
    return 0;
}

int ucg_set_linear_solver(int handle, int kind)
{
    std::lock_guard<std::mutex> lk(g_mu);
    NetHandle *h = get(handle);
    if (!h) return fail(UCG_E_HANDLE, "bad network handle");
    if (kind != UCG_LINSOL_SPARSE && kind != UCG_LINSOL_DENSE) return fail(UCG_E_ARG, "unknown linear solver");
    h->linear_solver = kind;
    return 0;
}

int ucg_network_info(int handle, int *nspec, int *neq, int *nreac)
{
    std::lock_guard<std::mutex> lk(g_mu);
    NetHandle *h = get(handle);
    if (!h) return fail(UCG_E_HANDLE, "bad network handle");
    if (nspec) *nspec = h->dev.nspec;
    if (neq) *neq = h->dev.neq;
    if (nreac) *nreac = h->dev.nreac;
    return 0;
}

int ucg_solve_batch_device(int handle, int n, const double *params, const double *y_init, double *y_out,
                           int max_rec, double *rec, int *n_rec, int *status, int *stats, double *t_start,
                           double *t_done, int lu_bits, int n_slots, void *stream)
{
    std::lock_guard<std::mutex> lk(g_mu);
    NetHandle *h = get(handle);
    if (!h) return fail(UCG_E_HANDLE, "bad network handle");
    if (n < 0) return fail(UCG_E_ARG, "n must not be negative");
    if (n == 0) return 0;
    const size_t un = (size_t)n, neq = (size_t)h->dev.neq, rw = 4 + (size_t)h->dev.nspec;
    if (!rec) max_rec = 0;
    if (max_rec < 0) return fail(UCG_E_ARG, "max_rec must not be negative");
    int rc;
    if ((rc = check_dev_buf(params, un * UCG_NPARAM * sizeof(double), "params")) ||
        (rc = check_dev_buf(y_init, un * neq * sizeof(double), "y_init", true)) ||
        (rc = check_dev_buf(y_out, un * neq * sizeof(double), "y_out")) ||
        (rc = check_dev_buf(rec, un * (size_t)max_rec * rw * sizeof(double), "rec", max_rec == 0)) ||
        (rc = check_dev_buf(n_rec, un * sizeof(int), "n_rec", true)) ||
        (rc = check_dev_buf(status, un * sizeof(int), "status")) ||
        (rc = check_dev_buf(stats, un * 8 * sizeof(int), "stats")) ||
        (rc = check_dev_buf(t_start, un * sizeof(double), "t_start", true)) ||
        (rc = check_dev_buf(t_done, un * sizeof(double), "t_done", true)))
        return rc;
    return launch_solve(*h, n, params, y_init, y_out, max_rec, rec, n_rec, status, stats, t_start, t_done,
                        lu_bits, n_slots, (hipStream_t)stream);
}

int ucg_solve_batch_device_noalloc(int handle, int n, const double *params, const double *y_init, double *y_out,
                                   int max_rec, double *rec, int *n_rec, int *status, int *stats, double *t_start,
                                   double *t_done, int lu_bits, int n_slots, void *stream)
{
    std::lock_guard<std::mutex> lk(g_mu);
    NetHandle *h = get(handle);
    if (!h) return fail(UCG_E_HANDLE, "bad network handle");
    if (n < 0) return fail(UCG_E_ARG, "n must not be negative");
    if (n == 0) return 0;
    if (lu_bits != 64 && lu_bits != 32) return fail(UCG_E_ARG, "lu_bits must be 64 or 32");
    if (n_slots < 0) return fail(UCG_E_ARG, "n_slots must not be negative");
    void *base = g_device >= 0 && g_device < (int)g_scratch.size() ? g_scratch[g_device] : nullptr;
    if (!base) return fail(UCG_E_HIP, "the fixed workspace of this GPU was not set up");
    const size_t stride = slot_layout(*h, lu_bits).stride;
    const size_t cap = (UCG_SCRATCH_BYTES - 256) / stride;   // slots the fixed workspace holds
    if (cap == 0)
        return fail(UCG_E_ARG, "one slot of this network needs " + std::to_string(stride) + " B, more than the " +
                                   std::to_string(UCG_SCRATCH_BYTES - 256) + " B fixed workspace");
    const int want = std::min(n_slots > 0 ? n_slots : 2 * g_num_cu, n);
    if (n_slots > 0 && (size_t)want > cap)
        return fail(UCG_E_ARG, std::to_string(want) + " slots of " + std::to_string(stride) +
                                   " B exceed the fixed workspace (at most " + std::to_string(cap) + " slots)");
    const int slots = (int)std::min((size_t)want, cap);
    const size_t un = (size_t)n, neq = (size_t)h->dev.neq, rw = 4 + (size_t)h->dev.nspec;
    if (!rec) max_rec = 0;
    if (max_rec < 0) return fail(UCG_E_ARG, "max_rec must not be negative");
    int rc;
    if ((rc = check_dev_buf(params, un * UCG_NPARAM * sizeof(double), "params")) ||
        (rc = check_dev_buf(y_init, un * neq * sizeof(double), "y_init", true)) ||
        (rc = check_dev_buf(y_out, un * neq * sizeof(double), "y_out")) ||
        (rc = check_dev_buf(rec, un * (size_t)max_rec * rw * sizeof(double), "rec", max_rec == 0)) ||
        (rc = check_dev_buf(n_rec, un * sizeof(int), "n_rec", true)) ||
        (rc = check_dev_buf(status, un * sizeof(int), "status")) ||
        (rc = check_dev_buf(stats, un * 8 * sizeof(int), "stats")) ||
        (rc = check_dev_buf(t_start, un * sizeof(double), "t_start", true)) ||
        (rc = check_dev_buf(t_done, un * sizeof(double), "t_done", true)))
        return rc;
    FixedWs fw;
    fw.counter = (int *)base;                        // reset by solve_init_kernel on the launch stream
    fw.ws = (unsigned char *)base + 256;
    return launch_solve(*h, n, params, y_init, y_out, max_rec, rec, n_rec, status, stats, t_start, t_done, lu_bits,
                        slots, (hipStream_t)stream, nullptr, IntervalArgs(), fw);
}

int ucg_solve_batch_profiled(int handle, int n, const double *params, double *y_out, int *status, int *stats,
                             double *timers, int lu_bits, int n_slots)
{
    std::lock_guard<std::mutex> lk(g_mu);
    NetHandle *h = get(handle);
    if (!h) return fail(UCG_E_HANDLE, "bad network handle");
    int rc = check_params(n, params);
    if (rc) return rc;
    const int neq = h->dev.neq;
    DevBuf dp, dy, dst, dstat, dtm;
    if ((rc = to_dev(dp, params, (size_t)n * UCG_NPARAM))) return rc;
    if ((rc = alloc_dev<double>(dy, (size_t)n * neq))) return rc;
    if ((rc = alloc_dev<int>(dst, (size_t)n))) return rc;
    if ((rc = alloc_dev<int>(dstat, (size_t)n * 8))) return rc;
    if ((rc = alloc_dev<double>(dtm, (size_t)n * 16))) return rc;
    rc = launch_solve(*h, n, (const double *)dp.p, nullptr, (double *)dy.p, 0, nullptr, nullptr, (int *)dst.p,
                      (int *)dstat.p, nullptr, nullptr, lu_bits, n_slots, nullptr, (double *)dtm.p);
    if (rc) return rc;
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(y_out, dy.p, (size_t)n * neq * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(status, dst.p, (size_t)n * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(stats, dstat.p, (size_t)n * 8 * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(timers, dtm.p, (size_t)n * 16 * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

static int solve_batch_host(int handle, int n, const double *params, const double *y_init, double *y_out, int max_rec,
                            double *rec, double *rates_rec, int *n_rec, int *status, int *stats, int lu_bits)
{
    std::lock_guard<std::mutex> lk(g_mu);
    NetHandle *h = get(handle);
    if (!h) return fail(UCG_E_HANDLE, "bad network handle");
    if (n <= 0 || !params || !y_out || !status || !stats) return fail(UCG_E_ARG, "bad arguments");
    int rc = check_params(n, params);
    if (rc) return rc;
    const int neq = h->dev.neq, ns = h->dev.nspec, nr = h->dev.nreac;
    if (!rec) max_rec = 0;
    DevBuf dp, dy0, dy, drec, drr, dnrec, dst, dstat;
    if ((rc = to_dev(dp, params, (size_t)n * UCG_NPARAM))) return rc;
    if (y_init && (rc = to_dev(dy0, y_init, (size_t)n * neq))) return rc;
    if ((rc = alloc_dev<double>(dy, (size_t)n * neq))) return rc;
    if (max_rec > 0 && (rc = alloc_dev<double>(drec, (size_t)n * max_rec * (4 + ns)))) return rc;
    if (max_rec > 0 && rates_rec && (rc = alloc_dev<double>(drr, (size_t)n * max_rec * nr))) return rc;
    if ((rc = alloc_dev<int>(dnrec, (size_t)n))) return rc;
    if ((rc = alloc_dev<int>(dst, (size_t)n))) return rc;
    if ((rc = alloc_dev<int>(dstat, (size_t)n * 8))) return rc;
    rc = launch_solve(*h, n, (const double *)dp.p, (const double *)dy0.p, (double *)dy.p, max_rec,
                      (double *)drec.p, (int *)dnrec.p, (int *)dst.p, (int *)dstat.p, nullptr, nullptr, lu_bits, 0,
                      nullptr, nullptr, IntervalArgs(), FixedWs(), (double *)drr.p);
    if (rc) return rc;
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(y_out, dy.p, (size_t)n * neq * sizeof(double), hipMemcpyDeviceToHost));
    if (max_rec > 0) HIPCHK(hipMemcpy(rec, drec.p, (size_t)n * max_rec * (4 + ns) * sizeof(double), hipMemcpyDeviceToHost));
    if (drr.p)
        HIPCHK(hipMemcpy(rates_rec, drr.p, (size_t)n * max_rec * nr * sizeof(double), hipMemcpyDeviceToHost));
    if (n_rec) HIPCHK(hipMemcpy(n_rec, dnrec.p, (size_t)n * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(status, dst.p, (size_t)n * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(stats, dstat.p, (size_t)n * 8 * sizeof(int), hipMemcpyDeviceToHost));
    return 0;
}

int ucg_solve_batch(int handle, int n, const double *params, const double *y_init, double *y_out, int max_rec,
                    double *rec, int *n_rec, int *status, int *stats, int lu_bits)
{
    return solve_batch_host(handle, n, params, y_init, y_out, max_rec, rec, nullptr, n_rec, status, stats, lu_bits);
}

int ucg_solve_batch_rates(int handle, int n, const double *params, const double *y_init, double *y_out, int max_rec,
                          double *rec, double *rates_rec, int *n_rec, int *status, int *stats, int lu_bits)
{
    if (!rec || max_rec <= 0 || !rates_rec) return fail(UCG_E_ARG, "rates output needs rec, max_rec > 0 and rates_rec");
    return solve_batch_host(handle, n, params, y_init, y_out, max_rec, rec, rates_rec, n_rec, status, stats, lu_bits);
}

int ucg_interval(int handle, int n, const double *params, const double *phys, const double *y,
                 const double *t_target, double *y_out, double *phys_out, int *status, int *stats, int lu_bits)
{
    std::lock_guard<std::mutex> lk(g_mu);
    NetHandle *h = get(handle);
    if (!h) return fail(UCG_E_HANDLE, "bad network handle");
    if (n <= 0 || !params || !phys || !y || !t_target || !y_out || !phys_out || !status || !stats)
        return fail(UCG_E_ARG, "bad arguments");
    int rc = check_params(n, params);
    if (rc) return rc;
    for (int p = 0; p < n; p++) {
        const double *ph = phys + (size_t)p * 8;
        if (!(ph[0] > 0) || !(ph[1] > 0) || !(ph[2] > 0) || !(ph[4] > 0) || !(ph[6] > 0) || !(ph[7] >= 0) ||
            !(t_target[p] >= ph[7]) || !std::isfinite(t_target[p]))
            return fail(UCG_E_PARAM, "invalid interval state for point " + std::to_string(p));
    }
    const size_t neq = h->dev.neq;
    DevBuf dp, dph, dy0, dtt, dy, dpo, dst, dstat;
    if ((rc = to_dev(dp, params, (size_t)n * UCG_NPARAM)) || (rc = to_dev(dph, phys, (size_t)n * 8)) ||
        (rc = to_dev(dy0, y, (size_t)n * neq)) || (rc = to_dev(dtt, t_target, (size_t)n)) ||
        (rc = alloc_dev<double>(dy, (size_t)n * neq)) || (rc = alloc_dev<double>(dpo, (size_t)n * 8)) ||
        (rc = alloc_dev<int>(dst, (size_t)n)) || (rc = alloc_dev<int>(dstat, (size_t)n * 8)))
        return rc;
    IntervalArgs iv;
    iv.phys_in = (const double *)dph.p;
    iv.t_target = (const double *)dtt.p;
    iv.phys_out = (double *)dpo.p;
    rc = launch_solve(*h, n, (const double *)dp.p, (const double *)dy0.p, (double *)dy.p, 0, nullptr, nullptr,
                      (int *)dst.p, (int *)dstat.p, nullptr, nullptr, lu_bits, 0, nullptr, nullptr, iv);
    if (rc) return rc;
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(y_out, dy.p, (size_t)n * neq * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(phys_out, dpo.p, (size_t)n * 8 * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(status, dst.p, (size_t)n * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(stats, dstat.p, (size_t)n * 8 * sizeof(int), hipMemcpyDeviceToHost));
    return 0;
}

int ucg_rates(int handle, int n, const double *params, const double *phys, const double *y, double *rate_out,
              double *y_out, double *scal_out)
{
    std::lock_guard<std::mutex> lk(g_mu);
    NetHandle *h = get(handle);
    if (!h) return fail(UCG_E_HANDLE, "bad network handle");
    if (!params || !phys || !y || !rate_out || !y_out || !scal_out) return fail(UCG_E_ARG, "null buffer");
    const ucg::DevNet &N = h->dev;
    ucg::PieceArgs A;
    DevBuf dp, dph, dy, dr, dyo, ds;
    int rc = piece_common(*h, n, params, phys, A, dp, dph);
    if (rc) return rc;
    if ((rc = to_dev(dy, y, (size_t)n * N.neq))) return rc;
    if ((rc = alloc_dev<double>(dr, (size_t)n * N.nreac))) return rc;
    if ((rc = alloc_dev<double>(dyo, (size_t)n * N.neq))) return rc;
    if ((rc = alloc_dev<double>(ds, (size_t)n * 8))) return rc;
    A.y = (const double *)dy.p;
    A.rate_out = (double *)dr.p;
    A.y_out = (double *)dyo.p;
    A.scal_out = (double *)ds.p;
    size_t lds = lds_bytes(N);
    HIPCHK(hipFuncSetAttribute((const void *)ucg::rates_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(ucg::rates_kernel, dim3(n), dim3(ucg::NT), lds, 0, h->d_dev, A);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(rate_out, dr.p, (size_t)n * N.nreac * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(y_out, dyo.p, (size_t)n * N.neq * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(scal_out, ds.p, (size_t)n * 8 * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int ucg_rhs(int handle, int n, const double *params, const double *phys, const double *rate, const double *y,
            double *ydot, double *rate_out)
{
    std::lock_guard<std::mutex> lk(g_mu);
    NetHandle *h = get(handle);
    if (!h) return fail(UCG_E_HANDLE, "bad network handle");
    if (!params || !phys || !rate || !y || !ydot) return fail(UCG_E_ARG, "null buffer");
    const ucg::DevNet &N = h->dev;
    ucg::PieceArgs A;
    DevBuf dp, dph, dy, drin, dyo, dro;
    int rc = piece_common(*h, n, params, phys, A, dp, dph);
    if (rc) return rc;
    if ((rc = to_dev(dy, y, (size_t)n * N.neq))) return rc;
    if ((rc = to_dev(drin, rate, (size_t)n * N.nreac))) return rc;
    if ((rc = alloc_dev<double>(dyo, (size_t)n * N.neq))) return rc;
    if (rate_out && (rc = alloc_dev<double>(dro, (size_t)n * N.nreac))) return rc;
    A.y = (const double *)dy.p;
    A.rate_in = (const double *)drin.p;
    A.y_out = (double *)dyo.p;
    A.rate_out = (double *)dro.p;
    size_t lds = lds_bytes(N);
    HIPCHK(hipFuncSetAttribute((const void *)ucg::rhs_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(ucg::rhs_kernel, dim3(n), dim3(ucg::NT), lds, 0, h->d_dev, A);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(ydot, dyo.p, (size_t)n * N.neq * sizeof(double), hipMemcpyDeviceToHost));
    if (rate_out) HIPCHK(hipMemcpy(rate_out, dro.p, (size_t)n * N.nreac * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int ucg_rhs_fluxes(int handle, int n, const double *params, const double *phys, const double *rate, const double *y,
                   double *ydot, double *flux)
{
    std::lock_guard<std::mutex> lk(g_mu);
    NetHandle *h = get(handle);
    if (!h) return fail(UCG_E_HANDLE, "bad network handle");
    if (!params || !phys || !rate || !y || !ydot || !flux) return fail(UCG_E_ARG, "null buffer");
    const ucg::DevNet &N = h->dev;
    ucg::PieceArgs A;
    DevBuf dp, dph, dy, drin, dyo, dfl;
    int rc = piece_common(*h, n, params, phys, A, dp, dph);
    if (rc) return rc;
    if ((rc = to_dev(dy, y, (size_t)n * N.neq))) return rc;
    if ((rc = to_dev(drin, rate, (size_t)n * N.nreac))) return rc;
    if ((rc = alloc_dev<double>(dyo, (size_t)n * N.neq))) return rc;
    if ((rc = alloc_dev<double>(dfl, (size_t)n * N.nreac))) return rc;
    A.y = (const double *)dy.p;
    A.rate_in = (const double *)drin.p;
    A.y_out = (double *)dyo.p;
    A.flux_out = (double *)dfl.p;
    size_t lds = lds_bytes(N);
    HIPCHK(hipFuncSetAttribute((const void *)ucg::rhs_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(ucg::rhs_kernel, dim3(n), dim3(ucg::NT), lds, 0, h->d_dev, A);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(ydot, dyo.p, (size_t)n * N.neq * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(flux, dfl.p, (size_t)n * N.nreac * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int ucg_jacobian(int handle, int n, const double *params, const double *phys, const double *rate, const double *y,
                 double *jac)
{
    std::lock_guard<std::mutex> lk(g_mu);
    NetHandle *h = get(handle);
    if (!h) return fail(UCG_E_HANDLE, "bad network handle");
    if (!params || !phys || !rate || !y || !jac) return fail(UCG_E_ARG, "null buffer");
    const ucg::DevNet &N = h->dev;
    ucg::PieceArgs A;
    DevBuf dp, dph, dy, drin, dj;
    int rc = piece_common(*h, n, params, phys, A, dp, dph);
    if (rc) return rc;
    if ((rc = to_dev(dy, y, (size_t)n * N.neq))) return rc;
    if ((rc = to_dev(drin, rate, (size_t)n * N.nreac))) return rc;
    if ((rc = alloc_dev<double>(dj, (size_t)n * N.neq * N.neq))) return rc;
    A.y = (const double *)dy.p;
    A.rate_in = (const double *)drin.p;
    A.jac_out = (double *)dj.p;
    size_t lds = lds_bytes(N);
    HIPCHK(hipFuncSetAttribute((const void *)ucg::jac_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(ucg::jac_kernel, dim3(n), dim3(ucg::NT), lds, 0, h->d_dev, A);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(jac, dj.p, (size_t)n * N.neq * N.neq * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int ucg_newton_solve(int handle, int n, const double *params, const double *phys, const double *y, double gamma,
                     const double *b, double *x, int lu_bits)
{
    std::lock_guard<std::mutex> lk(g_mu);
    NetHandle *h = get(handle);
    if (!h) return fail(UCG_E_HANDLE, "bad network handle");
    if (!params || !phys || !y || !b || !x) return fail(UCG_E_ARG, "null buffer");
    if (lu_bits != 64 && lu_bits != 32) return fail(UCG_E_ARG, "lu_bits must be 64 or 32");
    const ucg::DevNet &N = h->dev;
    ucg::PieceArgs A;
    DevBuf dp, dph, dy, db, dx;
    int rc = piece_common(*h, n, params, phys, A, dp, dph);
    if (rc) return rc;
    if ((rc = to_dev(dy, y, (size_t)n * N.neq))) return rc;
    if ((rc = to_dev(db, b, (size_t)n * N.neq))) return rc;
    if ((rc = alloc_dev<double>(dx, (size_t)n * N.neq))) return rc;
    A.y = (const double *)dy.p;
    A.scal_out = (double *)db.p;   // right-hand sides
    A.y_out = (double *)dx.p;
    A.gamma = gamma;
    A.sparse = h->linear_solver == UCG_LINSOL_SPARSE;
    size_t lds = lds_bytes(N);
    if (lu_bits == 64) {
        HIPCHK(hipFuncSetAttribute((const void *)ucg::newton_solve_kernel<double>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(ucg::newton_solve_kernel<double>, dim3(n), dim3(ucg::NT), lds, 0, h->d_dev, A);
    } else {
        HIPCHK(hipFuncSetAttribute((const void *)ucg::newton_solve_kernel<float>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(ucg::newton_solve_kernel<float>, dim3(n), dim3(ucg::NT), lds, 0, h->d_dev, A);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(x, dx.p, (size_t)n * N.neq * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

// ---------------------------------------------------------------- request queues

int ucg_queue_slot_bytes(int handle, int lu_bits, size_t *bytes)
{
    std::lock_guard<std::mutex> lk(g_mu);
    NetHandle *h = get(handle);
    if (!h) return fail(UCG_E_HANDLE, "bad network handle");
    if (lu_bits != 64 && lu_bits != 32) return fail(UCG_E_ARG, "lu_bits must be 64 or 32");
    if (!bytes) return fail(UCG_E_ARG, "null output");
    *bytes = slot_layout(*h, lu_bits).stride;
    return 0;
}

static int queue_open_locked(const int *handles, int nnet, int capacity, int n_slots, int lu_bits,
                             double idle_timeout_s, int *queue)
{
    if (!queue || !handles || nnet <= 0) return fail(UCG_E_ARG, "null output or no network");
    std::vector<NetHandle *> hs;
    for (int k = 0; k < nnet; k++) {
        NetHandle *h = get(handles[k]);
        if (!h) return fail(UCG_E_HANDLE, "bad network handle");
        if (h->linear_solver != g_nets[handles[0]].linear_solver)
            return fail(UCG_E_ARG, "the networks of a shared queue must use the same linear solver");
        hs.push_back(h);
    }
    if (capacity <= 0 || n_slots <= 0 || n_slots > g_num_cu)
        return fail(UCG_E_ARG, "capacity must be positive and n_slots in [1, number of CUs]");
    if (lu_bits != 64 && lu_bits != 32) return fail(UCG_E_ARG, "lu_bits must be 64 or 32");
    if (!(idle_timeout_s > 0.0) || idle_timeout_s > 3600.0) return fail(UCG_E_ARG, "idle timeout must be in (0, 3600] s");
    // one slot layout and one LDS size that fit every network of the queue
    size_t mat = 0, jac = 0, stride = 0, lds = 0;
    int neq_max = 0;
    for (NetHandle *h : hs) {
        const SlotLayout sl = slot_layout(*h, lu_bits);
        mat = std::max(mat, sl.mat);
        jac = std::max(jac, sl.jac);
        stride = std::max(stride, sl.stride - sl.mat - sl.jac);   // the rate coefficients
        lds = std::max(lds, lds_bytes(h->dev));
        neq_max = std::max(neq_max, h->dev.neq);
    }
    stride += mat + jac;
    int id = -1;
    for (int i = 0; i < (int)g_queues.size(); i++)
        if (!g_queues[i].used) {   // drained queues keep their slot until released
            id = i;
            break;
        }
    if (id < 0) {
        g_queues.emplace_back();
        id = (int)g_queues.size() - 1;
    }
    Queue &q = g_queues[id];
    q = Queue();
    q.used = true;
    q.handle = handles[0];
    q.handles.assign(handles, handles + nnet);
    q.capacity = capacity;
    q.n_slots = n_slots;
    q.lu_bits = lu_bits;
    q.y_stride = neq_max;
    NetHandle *h = hs[0];
    const unsigned fl = hipHostMallocMapped | hipHostMallocCoherent;
    hipError_t e = hipSuccess;
    auto hm = [&](void **p, size_t b) {
        if (e == hipSuccess) e = hipHostMalloc(p, b, fl);
        if (e == hipSuccess) memset(*p, 0, b);
    };
    hm((void **)&q.h_params, (size_t)capacity * UCG_NPARAM * sizeof(double));
    hm((void **)&q.h_ctrl, 2 * sizeof(int));
    hm((void **)&q.h_done, (size_t)capacity * sizeof(int));
    hm((void **)&q.h_y, (size_t)capacity * neq_max * sizeof(double));
    hm((void **)&q.h_status, (size_t)capacity * sizeof(int));
    hm((void **)&q.h_stats, (size_t)capacity * 8 * sizeof(int));
    if (nnet > 1) hm((void **)&q.h_net, (size_t)capacity * sizeof(int));
    hm((void **)&q.h_times, (size_t)capacity * 2 * sizeof(double));
    q.ws_bytes = stride * (size_t)n_slots;
    if (e == hipSuccess) e = hipMalloc(&q.ws, q.ws_bytes);
    if (e == hipSuccess) e = hipMalloc(&q.counter, sizeof(int));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&q.stream, hipStreamNonBlocking);
    if (e == hipSuccess && nnet > 1) {   // the descriptors' device addresses, in queue order
        std::vector<const ucg::DevNet *> dn;
        for (NetHandle *x : hs) dn.push_back(x->d_dev);
        e = hipMalloc(&q.d_nets, nnet * sizeof(const ucg::DevNet *));
        if (e == hipSuccess)
            e = hipMemcpyAsync(q.d_nets, dn.data(), nnet * sizeof(const ucg::DevNet *), hipMemcpyHostToDevice,
                               q.stream);
    }
    // stream-ordered clear: other queues' kernels may be running, and nothing here may wait
    // for them (a device-wide synchronisation would wait for kernels that only end on close)
    if (e == hipSuccess) e = hipMemsetAsync(q.counter, 0, sizeof(int), q.stream);
    ucg::SolveArgs A = {};
    void *d_params = nullptr, *d_ctrl = nullptr, *d_done = nullptr, *d_y = nullptr, *d_status = nullptr,
         *d_stats = nullptr, *d_net = nullptr, *d_times = nullptr;
    if (e == hipSuccess) e = hipHostGetDevicePointer(&d_params, q.h_params, 0);
    if (e == hipSuccess) e = hipHostGetDevicePointer(&d_ctrl, q.h_ctrl, 0);
    if (e == hipSuccess) e = hipHostGetDevicePointer(&d_done, q.h_done, 0);
    if (e == hipSuccess) e = hipHostGetDevicePointer(&d_y, q.h_y, 0);
    if (e == hipSuccess) e = hipHostGetDevicePointer(&d_status, q.h_status, 0);
    if (e == hipSuccess) e = hipHostGetDevicePointer(&d_stats, q.h_stats, 0);
    if (e == hipSuccess && q.h_net) e = hipHostGetDevicePointer(&d_net, q.h_net, 0);
    if (e == hipSuccess) e = hipHostGetDevicePointer(&d_times, q.h_times, 0);
    if (e != hipSuccess) {
        std::string msg = std::string("queue set-up: ") + hipGetErrorString(e);
        queue_release(q);
        return fail(UCG_E_HIP, msg);
    }
    A.n_points = capacity;
    A.params = (const double *)d_params;
    A.y_out = (double *)d_y;
    A.status = (int *)d_status;
    A.stats = (int *)d_stats;
    A.sparse = h->linear_solver == UCG_LINSOL_SPARSE ? 1 : 0;
    A.ws = q.ws;
    A.ws_stride = stride;
    A.ws_mat_bytes = mat;
    A.ws_jac_bytes = jac;
    A.counter = q.counter;
    A.cancel = h->cancel;   // ucg_cancel on the (first) network stops its queues too
    A.gen = ++h->gen;
    A.q_avail = (const int *)d_ctrl;
    A.q_closed = (const int *)d_ctrl + 1;
    A.q_done = (int *)d_done;
    A.q_idle_ticks = (long long)(idle_timeout_s * 1.0e8);
    A.t_start = (double *)d_times;
    A.t_done = (double *)d_times + capacity;
    A.q_nets = (const ucg::DevNet *const *)q.d_nets;
    A.q_net = (const int *)d_net;
    A.y_stride = neq_max;
    int rc = launch_kernel(h->dev, h->d_dev, A, n_slots, lu_bits, q.stream, lds);
    if (rc) {
        queue_release(q);
        return rc;
    }
    *queue = id;
    return 0;
}

int ucg_queue_open(int handle, int capacity, int n_slots, int lu_bits, double idle_timeout_s, int *queue)
{
    std::lock_guard<std::mutex> lk(g_mu);
    return queue_open_locked(&handle, 1, capacity, n_slots, lu_bits, idle_timeout_s, queue);
}

int ucg_queue_open_shared(const int *handles, int n_networks, int capacity, int n_slots, int lu_bits,
                          double idle_timeout_s, int *queue)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (n_networks > 64) return fail(UCG_E_ARG, "at most 64 networks per shared queue");
    return queue_open_locked(handles, n_networks, capacity, n_slots, lu_bits, idle_timeout_s, queue);
}

static Queue *get_queue(int queue)
{
    if (queue < 0 || queue >= (int)g_queues.size() || !g_queues[queue].used || g_queues[queue].drained) return nullptr;
    return &g_queues[queue];
}

static int queue_submit_locked(Queue *q, int network, const double *params, int *request)
{
    if (!params || !request) return fail(UCG_E_ARG, "null buffer");
    if (network < 0 || network >= (int)q->handles.size()) return fail(UCG_E_ARG, "bad network index for this queue");
    int rc = check_params(1, params);
    if (rc) return rc;
    if (q->next >= q->capacity) return fail(UCG_E_ARG, "queue capacity reached");
    if (__atomic_load_n(&q->h_ctrl[1], __ATOMIC_ACQUIRE)) return fail(UCG_E_ARG, "queue is closed");
    const int i = q->next++;
    memcpy(q->h_params + (size_t)i * UCG_NPARAM, params, UCG_NPARAM * sizeof(double));
    if (q->h_net) __atomic_store_n(&q->h_net[i], network, __ATOMIC_RELAXED);
    __atomic_store_n(&q->h_ctrl[0], i + 1, __ATOMIC_RELEASE);   // the parameters are visible first
    *request = i;
    return 0;
}

int ucg_queue_submit(int queue, const double *params, int *request)
{
    std::lock_guard<std::mutex> lk(g_mu);
    Queue *q = get_queue(queue);
    if (!q) return fail(UCG_E_HANDLE, "bad queue");
    if (q->handles.size() > 1) return fail(UCG_E_ARG, "a shared queue takes requests through ucg_queue_submit_to");
    return queue_submit_locked(q, 0, params, request);
}

int ucg_queue_submit_to(int queue, int network, const double *params, int *request)
{
    std::lock_guard<std::mutex> lk(g_mu);
    Queue *q = get_queue(queue);
    if (!q) return fail(UCG_E_HANDLE, "bad queue");
    return queue_submit_locked(q, network, params, request);
}

int ucg_queue_test(int queue, int request, int *done)
{
    std::lock_guard<std::mutex> lk(g_mu);
    Queue *q = get_queue(queue);
    if (!q) return fail(UCG_E_HANDLE, "bad queue");
    if (!done || request < 0 || request >= q->next) return fail(UCG_E_ARG, "bad request");
    *done = __atomic_load_n(&q->h_done[request], __ATOMIC_ACQUIRE);
    if (!*done && hipStreamQuery(q->stream) == hipSuccess)
        return fail(UCG_E_ARG, "the queue's kernel has exited (idle limit or cancel) before this request ran");
    return 0;
}

int ucg_queue_result(int queue, int request, double *y, int *status, int *stats)
{
    std::lock_guard<std::mutex> lk(g_mu);
    Queue *q = get_queue(queue);
    if (!q) return fail(UCG_E_HANDLE, "bad queue");
    if (request < 0 || request >= q->next) return fail(UCG_E_ARG, "bad request");
    if (!__atomic_load_n(&q->h_done[request], __ATOMIC_ACQUIRE)) return fail(UCG_E_ARG, "request not finished");
    const int net = q->h_net ? __atomic_load_n(&q->h_net[request], __ATOMIC_RELAXED) : 0;
    const int neq = g_nets[q->handles[net]].dev.neq;
    if (y) memcpy(y, q->h_y + (size_t)request * q->y_stride, (size_t)neq * sizeof(double));
    if (status) *status = q->h_status[request];
    if (stats) memcpy(stats, q->h_stats + (size_t)request * 8, 8 * sizeof(int));
    return 0;
}

int ucg_queue_times(int queue, int request, double *t_taken, double *t_done)
{
    std::lock_guard<std::mutex> lk(g_mu);
    Queue *q = get_queue(queue);
    if (!q) return fail(UCG_E_HANDLE, "bad queue");
    if (request < 0 || request >= q->next) return fail(UCG_E_ARG, "bad request");
    if (!__atomic_load_n(&q->h_done[request], __ATOMIC_ACQUIRE)) return fail(UCG_E_ARG, "request not finished");
    const double tick = 1.0e-8;   // the device wall clock runs at 100 MHz
    if (t_taken) *t_taken = q->h_times[request] * tick;
    if (t_done) *t_done = q->h_times[q->capacity + request] * tick;
    return 0;
}

int ucg_queue_close(int queue)
{
    hipStream_t st;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        Queue *q = get_queue(queue);
        if (!q) return fail(UCG_E_HANDLE, "bad queue");
        __atomic_store_n(&q->h_ctrl[1], 1, __ATOMIC_RELEASE);
        st = q->stream;
    }
    hipError_t e = hipStreamSynchronize(st);   // posted requests finish, idle workgroups leave
    std::lock_guard<std::mutex> lk(g_mu);
    Queue *q = get_queue(queue);
    if (!q) return fail(UCG_E_HANDLE, "queue closed concurrently");   // another thread's close won
    q->drained = true;
    // Freeing device or pinned memory may wait for every kernel on the device, including
    // other queues' kernels that only end when those queues close: release drained queues
    // once no queue is running.
    bool running = false;
    for (const Queue &o : g_queues) running |= o.used && !o.drained;
    if (!running)
        for (Queue &o : g_queues)
            if (o.used) queue_release(o);
    if (e != hipSuccess) return fail(UCG_E_HIP, std::string("queue drain: ") + hipGetErrorString(e));
    return 0;
}

}  // extern "C"
