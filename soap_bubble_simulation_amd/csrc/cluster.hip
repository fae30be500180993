// cluster.hip -- foam cluster dynamics (gfx950).
//
// FoamSimulator::step of foam_dynamics.rs:166-202 for a cluster of N bubbles:
// all-pairs interaction forces, buoyancy and drag, semi-implicit Euler with
// damping and the speed clamp, contact detection (update_connections,
// foam.rs:321-360) and coalescence (foam_dynamics.rs:205-228 with
// merge_bubbles, foam.rs:365-399).
//
//   cluster_pair_kernel       all-pairs forces: lane = bubble i of a 64-bubble
//                             tile, wave = one contiguous run of partners j
//                             (read with scalar loads, wave-uniform);
//                             FP32 pair force as the reference computes it,
//                             compensated FP32 sums, one partial per
//                             (block row, i)
//   cluster_integrate_kernel  buoyancy + the partials (in run order) +
//                             drag (compensated FP32), then the FP32
//                             integration; double-buffered
//                             state, so forces read only the old positions
//   cluster_contact_kernel    per (i, run): contacts j > i counted, the first
//                             coalescence candidate in (i, j) order
//   cluster_scan_kernel       one-block exclusive scan of the counts (i-major)
//   cluster_connect_kernel    BubbleConnection records written in (i, j) order
//
// Bubble state on the device: pr[i] = (x, y, z, radius), vt[i] = (vx, vy, vz,
// thickness_nm), FP32 like the reference's glam Vec3/f32 fields.  Ids, aspect
// ratio and refractive index stay on the host (ids are uploaded for the
// connection records).  Pair work is N^2: at N = 4096 one step is 16.8 M pair
// evaluations spread over 1024 blocks of 16 waves (runs of 16 partners).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "drainage.hpp"
#include "sf_common.hpp"

namespace sf {
namespace {

constexpr float kPiF = 3.14159265358979323846f;
constexpr float kAirDensity = 1.2f;     // foam_dynamics.rs:106
constexpr float kMaxSpeed = 0.1f;       // foam_dynamics.rs:182
constexpr float kAdhesion = 2.0f;       // foam_dynamics.rs:147
constexpr float kContactGap = 0.3f;     // foam.rs:342
constexpr int kTile = 64;               // bubbles i per block (one per lane)
constexpr int kRunsPerBlock = 4;        // partner runs per contact/connect block (one per wave)
constexpr int kPairRuns = 16;           // partner runs per pair-kernel block (one per wave)

struct Consts {
    float gdir[3];          // gravity.normalize_or_zero()
    float gmag;             // gravity.length()
    float density, vdw, k_rep, damping, drag_c;   // drag_c = 6 pi eta
    float surface_tension, coal_thr, dt;
    int n, run, nruns;      // bubbles, partners per run, runs (contact kernels)
    int prun, npruns;       // partners per run, runs (force kernel)
};

// compute_interaction_force (foam_dynamics.rs:116-154), force on a from the
// pair (a, b), added to (fx, fy, fz).  With the van der Waals product taken
// as vdw (r_a r_b), every scalar is symmetric in a, b and delta negates
// exactly, so swapping a and b gives exactly -force.  Every term needs
// distance < 3 (r_a + r_b): pairs clearly beyond (squared distance past
// 9 s^2 with a 1e-5 margin for rounding) return at once, the common case.
// The distance and 1 / distance use v_sqrt_f32 and v_rcp_f32 (1 ulp) where
// the reference divides, and overlap^1.5 is overlap sqrt(overlap) where it
// calls powf: the results move by ulps.
// Compensated (Kahan) FP32 accumulation for the folds of partial sums: s + c
// carries the running sum to about twice FP32 precision.
struct KSum {
    float s = 0.f, c = 0.f;
    __device__ inline void add(float x) {
#pragma clang fp contract(off)
        const float y = x - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ inline float value() const { return s - c; }
};

__device__ inline void pair_force_add(float4 a, float4 b, const Consts& c, float& fx, float& fy, float& fz) {
#pragma clang fp contract(off)
    const float dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    const float s = a.w + b.w;
    const float s3 = 3.0f * s;
    if (d2 > s3 * s3 * 1.00001f) return;
    const float dist = __builtin_amdgcn_sqrtf(d2);
    if (dist < 1e-6f) return;
    const float inv = __builtin_amdgcn_rcpf(dist);
    float m = 0.f;                                   // signed magnitude along b - a
    if (dist < s3 && dist > s) m += c.vdw * (a.w * b.w) * (inv * inv);
    const float overlap = s - dist;
    if (overlap > 0.f) m -= c.k_rep * (overlap * __builtin_amdgcn_sqrtf(overlap));   // overlap^1.5
    const float thr = s * 1.2f;
    if (dist < thr && dist > s * 0.95f) m += kAdhesion * (thr - dist);
    fx += dx * inv * m;
    fy += dy * inv * m;
    fz += dz * inv * m;
}

// Partner positions of a run are loaded once per 64, one per lane, and
// broadcast with v_readlane (the contact kernels).  Every lane of a wave
// takes part (lanes past the last bubble load a clamped index and write
// nothing), so each broadcast reads a loaded value.
__device__ inline float4 lane_bcast(const float4& v, int t) {
    return make_float4(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.x), t)),
                       __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.y), t)),
                       __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.z), t)),
                       __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.w), t)));
}

// One partial per (block row, bubble): the block's kPairRuns runs are folded
// in run order through LDS.  The partner index is wave-uniform, so partner
// records come through scalar loads and cost the vector ALU nothing.
__global__ __launch_bounds__(64 * kPairRuns) void cluster_pair_kernel(const float4* __restrict__ pr, Consts c,
                                                                     float* __restrict__ part) {
    __shared__ float s_f[kPairRuns][3][kTile];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int run = __builtin_amdgcn_readfirstlane(static_cast<int>(blockIdx.y) * kPairRuns + w);
    const int i = blockIdx.x * kTile + lane;
    // Within a run only partners closer than 3 (r_a + r_b) add a term, a
    // handful at cluster densities; plain FP32 sums, folded compensated.
    float fx = 0.f, fy = 0.f, fz = 0.f;
    if (run < c.npruns) {                // whole wave
        const float4 a = pr[i < c.n ? i : c.n - 1];
        const int j0 = run * c.prun, j1 = min(c.n, j0 + c.prun);
        for (int j = j0; j < j1; ++j) pair_force_add(a, pr[j], c, fx, fy, fz);
    }
    s_f[w][0][lane] = fx;
    s_f[w][1][lane] = fy;
    s_f[w][2][lane] = fz;
    __syncthreads();
    if (w == 0 && i < c.n) {
        float* dst = part + (static_cast<size_t>(blockIdx.y) * c.n + i) * 3;
        for (int d = 0; d < 3; ++d) {
            KSum k;
            for (int r = 0; r < kPairRuns; ++r) k.add(s_f[r][d][lane]);
            dst[d] = k.value();
        }
    }
}

// compute_buoyancy (foam_dynamics.rs:96-111) magnitude, FP32.
__device__ inline float buoyancy_mag(float r, float t_nm, const Consts& c) {
#pragma clang fp contract(off)
    const float shell = 4.0f * kPiF * (r * r) * (t_nm * 1e-9f);
    const float film_mass = shell * c.density;
    const float vol = (4.0f / 3.0f) * kPiF * (r * r * r);
    const float net = film_mass * c.gmag - vol * kAirDensity * c.gmag;
    return fmaxf(net, 0.f);
}

__global__ __launch_bounds__(256) void cluster_integrate_kernel(const float4* __restrict__ pr,
                                                                const float4* __restrict__ vt,
                                                                const float* __restrict__ part, Consts c,
                                                                float4* __restrict__ pr_out,
                                                                float4* __restrict__ vt_out,
                                                                float* __restrict__ force_out) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n) return;
    const float4 p = pr[i];
    const float4 v = vt[i];
    const float bm = buoyancy_mag(p.w, v.w, c);
    const float dc = c.drag_c * p.w;     // 6 pi eta R (foam_dynamics.rs:159)
    KSum fx, fy, fz;
    fx.add(c.gdir[0] * bm);
    fy.add(c.gdir[1] * bm);
    fz.add(c.gdir[2] * bm);
    const int rows = (c.npruns + kPairRuns - 1) / kPairRuns;
    // Partials in row order; eight rows' loads are issued before their adds
    // so that the loads overlap (the order of the sum is unchanged).
    int r = 0;
    for (; r + 8 <= rows; r += 8) {
        float q[8][3];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float* src = part + (static_cast<size_t>(r + k) * c.n + i) * 3;
            q[k][0] = src[0];
            q[k][1] = src[1];
            q[k][2] = src[2];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            fx.add(q[k][0]);
            fy.add(q[k][1]);
            fz.add(q[k][2]);
        }
    }
    for (; r < rows; ++r) {
        const float* src = part + (static_cast<size_t>(r) * c.n + i) * 3;
        fx.add(src[0]); fy.add(src[1]); fz.add(src[2]);
    }
    fx.add(-v.x * dc);
    fy.add(-v.y * dc);
    fz.add(-v.z * dc);
    const float Fx = fx.value(), Fy = fy.value(), Fz = fz.value();
    if (force_out) {
        force_out[3 * i + 0] = Fx;
        force_out[3 * i + 1] = Fy;
        force_out[3 * i + 2] = Fz;
    }
    if (!pr_out) return;
    // Semi-implicit Euler (foam_dynamics.rs:171-192).
    const float shell = 4.0f * kPiF * (p.w * p.w) * (v.w * 1e-9f);
    const float mass = fmaxf(shell * c.density, 1e-9f);
    float vx = v.x + Fx / mass * c.dt, vy = v.y + Fy / mass * c.dt, vz = v.z + Fz / mass * c.dt;
    vx *= c.damping; vy *= c.damping; vz *= c.damping;
    const float speed = sqrtf(vx * vx + vy * vy + vz * vz);
    if (speed > kMaxSpeed) {
        const float k = kMaxSpeed / speed;
        vx *= k; vy *= k; vz *= k;
    }
    pr_out[i] = make_float4(p.x + vx * c.dt, p.y + vy * c.dt, p.z + vz * c.dt, p.w);
    vt_out[i] = make_float4(vx, vy, vz, v.w);
}

// Contact of the pair (i, j), i < j: overlap_distance > -0.3 r_i
// (foam.rs:80-83, 338-344).
__device__ inline bool in_contact(float4 a, float4 b) {
#pragma clang fp contract(off)
    const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    // The test holds only for distance < 1.3 r_a + r_b; clearly farther
    // pairs (1e-5 margin on the square) skip the square root.
    const float reach = 1.3f * a.w + b.w;
    if (d2 > reach * reach * 1.00001f) return false;
    const float dist = sqrtf(d2);
    return (a.w + b.w - dist) > -a.w * kContactGap;
}

__global__ __launch_bounds__(256) void cluster_contact_kernel(const float4* __restrict__ pr,
                                                              const float4* __restrict__ vt, Consts c,
                                                              int want_candidate, int* __restrict__ count,
                                                              long long* __restrict__ cand) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const int run = blockIdx.y * kRunsPerBlock + (threadIdx.x >> 6);
    const int i = blockIdx.x * kTile + lane;
    if (run >= c.nruns) return;          // whole wave
    const int ic = i < c.n ? i : c.n - 1;
    const float4 a = pr[ic];
    const float4 va = vt[ic];
    const int j0 = run * c.run, j1 = min(c.n, j0 + c.run);
    int k = 0;
    long long first = -1;
    for (int jc = j0; jc < j1; jc += 64) {
        const float4 mine = pr[min(jc + lane, c.n - 1)];
        const int cnt = min(64, j1 - jc);
        for (int t = 0; t < cnt; ++t) {
            const int j = jc + t;
            const float4 b = lane_bcast(mine, t);
            if (j <= i || !in_contact(a, b)) continue;
            ++k;
            if (want_candidate && first < 0) {
                // Weber number of the connection (foam_dynamics.rs:213-219).
                const float dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
                const float dist = sqrtf(dx * dx + dy * dy + dz * dz);
                const float nx = dx / dist, ny = dy / dist, nz = dz / dist;
                const float4 vb = vt[j];
                const float rel = (vb.x - va.x) * nx + (vb.y - va.y) * ny + (vb.z - va.z) * nz;
                const float weber = c.density * (rel * rel) * a.w / c.surface_tension;
                if (weber > c.coal_thr) first = static_cast<long long>(i) * c.n + j;
            }
        }
    }
    if (i >= c.n) return;
    count[static_cast<size_t>(i) * c.nruns + run] = k;
    if (want_candidate) cand[static_cast<size_t>(i) * c.nruns + run] = first;
}

// Exclusive scan of count[m] into offs[m], total in offs[m]; first
// coalescence candidate (smallest key >= 0, else -1) into cand_out.
__global__ __launch_bounds__(1024) void cluster_scan_kernel(const int* __restrict__ count, size_t m,
                                                            int* __restrict__ offs,
                                                            const long long* __restrict__ cand,
                                                            long long* __restrict__ cand_out) {
    __shared__ int s_sum[1024];
    __shared__ long long s_min[1024];
    const int t = threadIdx.x;
    const size_t per = (m + 1023) / 1024;
    const size_t lo = std::min(m, t * per), hi = std::min(m, lo + per);
    int local = 0;
    long long best = -1;
    for (size_t k = lo; k < hi; ++k) {
        local += count[k];
        if (cand) {
            const long long x = cand[k];
            if (x >= 0 && (best < 0 || x < best)) best = x;
        }
    }
    s_sum[t] = local;
    s_min[t] = best;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const int add = t >= d ? s_sum[t - d] : 0;
        long long other = t >= d ? s_min[t - d] : -1;
        __syncthreads();
        s_sum[t] += add;
        if (other >= 0 && (s_min[t] < 0 || other < s_min[t])) s_min[t] = other;
        __syncthreads();
    }
    int run = s_sum[t] - local;    // exclusive prefix of this thread's chunk
    for (size_t k = lo; k < hi; ++k) {
        offs[k] = run;
        run += count[k];
    }
    if (t == 1023) {
        offs[m] = s_sum[1023];
        if (cand_out) *cand_out = s_min[1023];
    }
}

struct ConnRec {
    uint32_t a, b;
    float point[3], normal[3];
    float wall_radius, wall_thickness_nm;
};
static_assert(sizeof(ConnRec) == sizeof(sf_cluster_connection), "connection record layout");

__global__ __launch_bounds__(256) void cluster_connect_kernel(const float4* __restrict__ pr,
                                                              const float4* __restrict__ vt,
                                                              const uint32_t* __restrict__ ids, Consts c,
                                                              const int* __restrict__ offs,
                                                              ConnRec* __restrict__ out) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const int run = blockIdx.y * kRunsPerBlock + (threadIdx.x >> 6);
    const int i = blockIdx.x * kTile + lane;
    if (run >= c.nruns) return;          // whole wave
    const int ic = i < c.n ? i : c.n - 1;
    const float4 a = pr[ic];
    const float ta = vt[ic].w;
    const float pa = 4.0f * c.surface_tension / a.w;     // pressure_difference (foam.rs:72-74)
    const int j0 = run * c.run, j1 = min(c.n, j0 + c.run);
    int k = i < c.n ? offs[static_cast<size_t>(i) * c.nruns + run] : 0;
    for (int jc = j0; jc < j1; jc += 64) {
        const float4 mine = pr[min(jc + lane, c.n - 1)];
        const int cnt = min(64, j1 - jc);
        for (int t = 0; t < cnt; ++t) {
            const int j = jc + t;
            const float4 b = lane_bcast(mine, t);
            if (i >= c.n || j <= i || !in_contact(a, b)) continue;
            // BubbleConnection::new (foam.rs:106-139).
            const float dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
            const float dist = sqrtf(dx * dx + dy * dy + dz * dz);
            const float tt = a.w / (a.w + b.w);
            ConnRec r;
            r.a = ids[i];
            r.b = ids[j];
            r.point[0] = a.x + dx * tt; r.point[1] = a.y + dy * tt; r.point[2] = a.z + dz * tt;
            r.normal[0] = dx / dist; r.normal[1] = dy / dist; r.normal[2] = dz / dist;
            const float dp = pa - 4.0f * c.surface_tension / b.w;
            r.wall_radius = fabsf(dp) > 1e-6f ? 2.0f * c.surface_tension / dp : __builtin_inff();
            r.wall_thickness_nm = (ta + vt[j].w) / 2.0f * 0.8f;
            out[k++] = r;
        }
    }
}

template <class T>
T* dnew(size_t count) {
    void* p = nullptr;
    if (count == 0) count = 1;
    const hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) {
        set_error("hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
        throw Error{SF_ERR_NOMEM};
    }
    return static_cast<T*>(p);
}

void dfree(void* p) {
    if (p) (void)hipFree(p);
}

}  // namespace

// Host object of sf_cluster.  The device holds the kinematic state; ids,
// aspect ratios and refractive indices live on the host.  The connection
// list is built on demand from the current state (it depends on nothing
// else), except across a merge, where the list of the step (before the
// merge, less the two merged bubbles) is kept as the reference keeps it.
struct Cluster {
    sf_cluster_params p{};
    int device = 0;
    hipStream_t stream = nullptr;
    int n = 0, cap = 0;
    float4* pr[2] = {nullptr, nullptr};
    float4* vt[2] = {nullptr, nullptr};
    int cur = 0;
    uint32_t* ids_dev = nullptr;
    std::vector<uint32_t> ids;
    std::vector<float> aspect, refr;
    uint32_t next_id = 0;
    float* part = nullptr;
    size_t part_cap = 0;
    int* count = nullptr;
    int* offs = nullptr;
    long long* cand = nullptr;
    long long* cand_out = nullptr;
    size_t cnt_cap = 0;
    ConnRec* conn_dev = nullptr;
    size_t conn_cap = 0;
    // Host connection list and neighbour lists (valid when conn_valid).
    bool conn_valid = false;
    std::vector<sf_cluster_connection> conns;
    std::vector<std::vector<uint32_t>> neighbors;
    KernelTimer pair_timer;
    // kGraphSteps steps (force + integration launches) captured in a HIP
    // graph for long runs without coalescence; valid for the key it was
    // captured with (the ping-pong phase returns after an even step count).
    static constexpr int kGraphSteps = 16;
    hipGraphExec_t graph = nullptr;
    struct GraphKey {
        sf_cluster_params p;
        float dt;
        int n, cur;
        const void* buf[5];
    } gkey{};

    ~Cluster() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (graph) (void)hipGraphExecDestroy(graph);
        pair_timer.destroy();
        for (int k = 0; k < 2; ++k) {
            dfree(pr[k]);
            dfree(vt[k]);
        }
        dfree(ids_dev); dfree(part); dfree(count); dfree(offs); dfree(cand); dfree(cand_out); dfree(conn_dev);
        if (stream) (void)hipStreamDestroy(stream);
    }

    Consts consts(float dt) const {
        Consts c{};
        // glam length (f32) and normalize_or_zero
        const float gl = sqrtf(p.gravity[0] * p.gravity[0] + p.gravity[1] * p.gravity[1] +
                               p.gravity[2] * p.gravity[2]);
        c.gmag = gl;
        const float rcp = 1.f / gl;
        for (int k = 0; k < 3; ++k) c.gdir[k] = (std::isfinite(rcp) && rcp > 0.f) ? p.gravity[k] * rcp : 0.f;
        c.density = p.density;
        c.vdw = p.van_der_waals_strength;
        c.k_rep = p.repulsion_stiffness;
        c.damping = p.damping;
        c.drag_c = 6.0f * kPiF * p.viscosity;
        c.surface_tension = p.surface_tension;
        c.coal_thr = p.coalescence_threshold;
        c.dt = dt;
        c.n = n;
        c.run = std::max(kTile, (n + 63) / 64);
        c.nruns = n > 0 ? (n + c.run - 1) / c.run : 0;
        // Short force runs: enough waves to hide latency (16 per CU at
        // n = 4096), folded 16 to a block.
        c.prun = std::max(16, (n + 255) / 256);
        c.npruns = n > 0 ? (n + c.prun - 1) / c.prun : 0;
        return c;
    }

    void reserve(int want) {
        if (want <= cap) return;
        const int nc = std::max(want, std::max(64, cap * 2));
        float4* npr[2] = {dnew<float4>(nc), dnew<float4>(nc)};
        float4* nvt[2] = {dnew<float4>(nc), dnew<float4>(nc)};
        uint32_t* nid = dnew<uint32_t>(nc);
        if (n > 0) {
            SF_HIP(hipMemcpyAsync(npr[0], pr[cur], n * sizeof(float4), hipMemcpyDeviceToDevice, stream));
            SF_HIP(hipMemcpyAsync(nvt[0], vt[cur], n * sizeof(float4), hipMemcpyDeviceToDevice, stream));
            SF_HIP(hipMemcpyAsync(nid, ids_dev, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream));
            SF_HIP(hipStreamSynchronize(stream));
        }
        for (int k = 0; k < 2; ++k) {
            dfree(pr[k]);
            dfree(vt[k]);
            pr[k] = npr[k];
            vt[k] = nvt[k];
        }
        dfree(ids_dev);
        ids_dev = nid;
        cur = 0;
        cap = nc;
    }

    void ensure_scratch(const Consts& c) {
        const size_t np = static_cast<size_t>((c.npruns + kPairRuns - 1) / kPairRuns) * n * 3;
        if (np > part_cap) {
            dfree(part);
            part = dnew<float>(np);
            part_cap = np;
        }
        const size_t m = static_cast<size_t>(c.nruns) * n;
        if (m + 1 > cnt_cap) {
            dfree(count); dfree(offs); dfree(cand);
            count = dnew<int>(m + 1);
            offs = dnew<int>(m + 1);
            cand = dnew<long long>(m + 1);
            cnt_cap = m + 1;
        }
        if (!cand_out) cand_out = dnew<long long>(1);
    }

    dim3 pair_grid(const Consts& c) const {
        return dim3((n + kTile - 1) / kTile, (c.nruns + kRunsPerBlock - 1) / kRunsPerBlock);
    }

    dim3 force_grid(const Consts& c) const {
        return dim3((n + kTile - 1) / kTile, (c.npruns + kPairRuns - 1) / kPairRuns);
    }

    // compute_forces into force_out (device, [n][3]) and, when out state is
    // given, the integrated state into the other buffers.
    void forces_and_integrate(float dt, float* force_dev, bool integrate) {
        if (n == 0) return;
        const Consts c = consts(dt);
        ensure_scratch(c);
        if (pair_timer.enabled) {
            auto& ev = pair_timer.next(stream);
            hipExtLaunchKernelGGL(cluster_pair_kernel, force_grid(c), dim3(64 * kPairRuns), 0, stream, ev.first,
                                  ev.second, 0, pr[cur], c, part);
        } else {
            hipLaunchKernelGGL(cluster_pair_kernel, force_grid(c), dim3(64 * kPairRuns), 0, stream, pr[cur], c, part);
        }
        SF_HIP(hipGetLastError());
        hipLaunchKernelGGL(cluster_integrate_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, pr[cur], vt[cur],
                           part, c, integrate ? pr[1 - cur] : nullptr, integrate ? vt[1 - cur] : nullptr,
                           force_dev);
        SF_HIP(hipGetLastError());
        if (integrate) cur = 1 - cur;
    }

    // Contact counts, their scan, and (when asked) the first coalescence
    // candidate; returns the connection count (synchronises).
    int contacts(bool candidate, long long* cand_key) {
        if (n == 0) {
            if (cand_key) *cand_key = -1;
            return 0;
        }
        const Consts c = consts(0.f);
        ensure_scratch(c);
        hipLaunchKernelGGL(cluster_contact_kernel, pair_grid(c), dim3(256), 0, stream, pr[cur], vt[cur], c,
                           candidate ? 1 : 0, count, cand);
        SF_HIP(hipGetLastError());
        const size_t m = static_cast<size_t>(c.nruns) * n;
        hipLaunchKernelGGL(cluster_scan_kernel, dim3(1), dim3(1024), 0, stream, count, m, offs,
                           candidate ? cand : nullptr, candidate ? cand_out : nullptr);
        SF_HIP(hipGetLastError());
        int total = 0;
        SF_HIP(hipMemcpyAsync(&total, offs + m, sizeof(int), hipMemcpyDeviceToHost, stream));
        long long key = -1;
        if (candidate) SF_HIP(hipMemcpyAsync(&key, cand_out, sizeof(long long), hipMemcpyDeviceToHost, stream));
        SF_HIP(hipStreamSynchronize(stream));
        if (cand_key) *cand_key = key;
        return total;
    }

    // update_connections (foam.rs:321-360) on the current state, after
    // contacts() has left the scanned offsets.
    void build_connections(int total) {
        const Consts c = consts(0.f);
        if (static_cast<size_t>(total) > conn_cap) {
            dfree(conn_dev);
            conn_cap = std::max<size_t>(total, 2 * conn_cap);
            conn_dev = dnew<ConnRec>(conn_cap);
        }
        conns.resize(total);
        if (total > 0) {
            hipLaunchKernelGGL(cluster_connect_kernel, pair_grid(c), dim3(256), 0, stream, pr[cur], vt[cur],
                               ids_dev, c, offs, conn_dev);
            SF_HIP(hipGetLastError());
            SF_HIP(hipMemcpyAsync(conns.data(), conn_dev, total * sizeof(ConnRec), hipMemcpyDeviceToHost, stream));
            SF_HIP(hipStreamSynchronize(stream));
        }
        // Neighbour lists in connection order (foam.rs:351-359).
        std::vector<int> index_of;
        neighbors.assign(n, {});
        std::vector<std::pair<uint32_t, int>> order(n);
        for (int k = 0; k < n; ++k) order[k] = {ids[k], k};
        std::sort(order.begin(), order.end());
        auto find = [&](uint32_t id) {
            auto it = std::lower_bound(order.begin(), order.end(), std::make_pair(id, -1));
            return (it != order.end() && it->first == id) ? it->second : -1;
        };
        for (const auto& r : conns) {
            const int ia = find(r.bubble_a), ib = find(r.bubble_b);
            if (ia >= 0) neighbors[ia].push_back(r.bubble_b);
            if (ib >= 0) neighbors[ib].push_back(r.bubble_a);
        }
        conn_valid = true;
    }

    void ensure_connections() {
        if (conn_valid) return;
        build_connections(contacts(false, nullptr));
    }

    void download(std::vector<float4>& hp, std::vector<float4>& hv) {
        hp.resize(n);
        hv.resize(n);
        if (n > 0) {
            SF_HIP(hipMemcpyAsync(hp.data(), pr[cur], n * sizeof(float4), hipMemcpyDeviceToHost, stream));
            SF_HIP(hipMemcpyAsync(hv.data(), vt[cur], n * sizeof(float4), hipMemcpyDeviceToHost, stream));
            SF_HIP(hipStreamSynchronize(stream));
        }
    }

    void upload(const std::vector<float4>& hp, const std::vector<float4>& hv) {
        reserve(static_cast<int>(hp.size()));
        n = static_cast<int>(hp.size());
        if (n > 0) {
            SF_HIP(hipMemcpyAsync(pr[cur], hp.data(), n * sizeof(float4), hipMemcpyHostToDevice, stream));
            SF_HIP(hipMemcpyAsync(vt[cur], hv.data(), n * sizeof(float4), hipMemcpyHostToDevice, stream));
            SF_HIP(hipMemcpyAsync(ids_dev, ids.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
            SF_HIP(hipStreamSynchronize(stream));
        }
    }

    int index_of(uint32_t id) const {
        for (int k = 0; k < n; ++k)
            if (ids[k] == id) return k;
        return -1;
    }

    // remove_bubble (foam.rs:249-259) on host copies of the state.
    bool remove_host(uint32_t id, std::vector<float4>& hp, std::vector<float4>& hv) {
        const int k = index_of(id);
        if (k < 0) return false;
        hp.erase(hp.begin() + k);
        hv.erase(hv.begin() + k);
        ids.erase(ids.begin() + k);
        aspect.erase(aspect.begin() + k);
        refr.erase(refr.begin() + k);
        if (conn_valid) {
            if (k < static_cast<int>(neighbors.size())) neighbors.erase(neighbors.begin() + k);
            conns.erase(std::remove_if(conns.begin(), conns.end(),
                                       [&](const sf_cluster_connection& r) {
                                           return r.bubble_a == id || r.bubble_b == id;
                                       }),
                        conns.end());
        }
        --n;
        return true;
    }

    // add_bubble (foam.rs:235-244) with Bubble::new defaults (foam.rs:52-63).
    uint32_t add_host(const float pos[3], float radius, std::vector<float4>& hp, std::vector<float4>& hv) {
        const uint32_t id = next_id++;
        hp.push_back(make_float4(pos[0], pos[1], pos[2], radius));
        hv.push_back(make_float4(0.f, 0.f, 0.f, 500.f));
        ids.push_back(id);
        aspect.push_back(1.f);
        refr.push_back(1.33f);
        if (conn_valid) neighbors.emplace_back();
        ++n;
        return id;
    }

    // merge_bubbles (foam.rs:365-399), FP32 as the reference.
    uint32_t merge(uint32_t id_a, uint32_t id_b) {
#pragma clang fp contract(off)
        std::vector<float4> hp, hv;
        download(hp, hv);
        const int ia = index_of(id_a), ib = index_of(id_b);
        SF_REQUIRE(ia >= 0 && ib >= 0 && ia != ib, SF_ERR_INVALID, "merge: unknown or equal bubble ids");
        const float4 pa = hp[ia], pb = hp[ib], va = hv[ia], vb = hv[ib];
        const float vol_a = (4.0f / 3.0f) * kPiF * (pa.w * pa.w * pa.w);
        const float vol_b = (4.0f / 3.0f) * kPiF * (pb.w * pb.w * pb.w);
        const float vol = vol_a + vol_b;
        const float r = powf(vol * 3.0f / (4.0f * kPiF), 1.0f / 3.0f);
        const float pos[3] = {(pa.x * vol_a + pb.x * vol_b) / vol, (pa.y * vol_a + pb.y * vol_b) / vol,
                              (pa.z * vol_a + pb.z * vol_b) / vol};
        const float vel[3] = {(va.x * vol_a + vb.x * vol_b) / vol, (va.y * vol_a + vb.y * vol_b) / vol,
                              (va.z * vol_a + vb.z * vol_b) / vol};
        const float area_a = 4.0f * kPiF * (pa.w * pa.w), area_b = 4.0f * kPiF * (pb.w * pb.w);
        const float th = (va.w * area_a + vb.w * area_b) / (area_a + area_b);
        remove_host(id_a, hp, hv);
        remove_host(id_b, hp, hv);
        const uint32_t id = add_host(pos, r, hp, hv);
        hv.back() = make_float4(vel[0], vel[1], vel[2], th);
        upload(hp, hv);
        return id;
    }

    GraphKey key_now(float dt) const {
        GraphKey k;
        std::memset(&k, 0, sizeof(k));
        k.p = p;
        k.dt = dt;
        k.n = n;
        k.cur = cur;
        k.buf[0] = pr[0]; k.buf[1] = pr[1]; k.buf[2] = vt[0]; k.buf[3] = vt[1]; k.buf[4] = part;
        return k;
    }

    void capture(float dt) {
        if (graph) {
            (void)hipGraphExecDestroy(graph);
            graph = nullptr;
        }
        ensure_scratch(consts(dt));
        const GraphKey k = key_now(dt);
        hipGraph_t gr = nullptr;
        SF_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        for (int s = 0; s < kGraphSteps; ++s) forces_and_integrate(dt, nullptr, true);
        hipError_t e = hipStreamEndCapture(stream, &gr);
        if (e == hipSuccess) e = hipGraphInstantiate(&graph, gr, nullptr, nullptr, 0);
        if (gr) (void)hipGraphDestroy(gr);
        SF_HIP(e);
        gkey = k;
    }

    // n_steps FoamSimulator steps; without coalescence, runs of kGraphSteps
    // replay a captured graph.
    void step_many(int n_steps, float dt) {
        int done = 0;
        if (!p.coalescence_enabled && !pair_timer.enabled && n > 0) {
            while (n_steps - done >= kGraphSteps) {
                ensure_scratch(consts(dt));
                const GraphKey k = key_now(dt);
                if (!graph || std::memcmp(&k, &gkey, sizeof(k)) != 0) capture(dt);
                SF_HIP(hipGraphLaunch(graph, stream));
                conn_valid = false;
                done += kGraphSteps;
            }
        }
        for (; done < n_steps; ++done) step(dt);
    }

    // FoamSimulator::step (foam_dynamics.rs:166-202).
    void step(float dt) {
        forces_and_integrate(dt, nullptr, true);
        conn_valid = false;
        if (p.coalescence_enabled) {
            long long key = -1;
            const int total = contacts(true, &key);
            if (key >= 0) {
                build_connections(total);
                const int i = static_cast<int>(key / n), j = static_cast<int>(key % n);
                merge(ids[i], ids[j]);
            }
        }
    }
};

}  // namespace sf

using sf::Cluster;

static Cluster* CL(sf_cluster* c) { return reinterpret_cast<Cluster*>(c); }

extern "C" int sf_cluster_params_default(sf_cluster_params* p) {
    return sf::guard([&] {
        SF_REQUIRE(p != nullptr, SF_ERR_INVALID, "p is NULL");
        // FoamSimulator::new (foam_dynamics.rs:37-53).
        p->surface_tension = 0.025f;
        p->viscosity = 0.001f;
        p->density = 1000.0f;
        p->gravity[0] = 0.f;
        p->gravity[1] = -9.81f;
        p->gravity[2] = 0.f;
        p->van_der_waals_strength = 0.001f;
        p->repulsion_stiffness = 5.0f;
        p->damping = 0.8f;
        p->coalescence_enabled = 0;
        p->coalescence_threshold = 0.5f;
    });
}

static void check_bubbles(int32_t n, const sf_cluster_bubble* b) {
    SF_REQUIRE(n >= 0 && (n == 0 || b != nullptr), SF_ERR_INVALID, "bad bubble array");
    for (int k = 0; k < n; ++k) {
        SF_REQUIRE(b[k].radius > 0.f && std::isfinite(b[k].radius), SF_ERR_INVALID, "bubble %d: radius must be > 0", k);
        for (int d = 0; d < 3; ++d)
            SF_REQUIRE(std::isfinite(b[k].position[d]) && std::isfinite(b[k].velocity[d]), SF_ERR_INVALID,
                       "bubble %d: position and velocity must be finite", k);
    }
}

extern "C" int sf_cluster_create(const sf_cluster_params* params, int32_t n, const sf_cluster_bubble* bubbles,
                                 sf_cluster** out) {
    return sf::guard([&] {
        SF_REQUIRE(params != nullptr && out != nullptr, SF_ERR_INVALID, "NULL argument");
        check_bubbles(n, bubbles);
        sf::require_device();
        auto c = std::make_unique<Cluster>();
        c->p = *params;
        SF_HIP(hipGetDevice(&c->device));
        SF_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        std::vector<float4> hp(n), hv(n);
        for (int k = 0; k < n; ++k) {
            const sf_cluster_bubble& b = bubbles[k];
            hp[k] = make_float4(b.position[0], b.position[1], b.position[2], b.radius);
            hv[k] = make_float4(b.velocity[0], b.velocity[1], b.velocity[2], b.thickness_nm);
            c->ids.push_back(b.id);
            c->aspect.push_back(b.aspect_ratio);
            c->refr.push_back(b.refractive_index);
            c->next_id = std::max(c->next_id, b.id + 1);
        }
        c->upload(hp, hv);
        *out = reinterpret_cast<sf_cluster*>(c.release());
    });
}

extern "C" void sf_cluster_destroy(sf_cluster* c) { delete CL(c); }

extern "C" int sf_cluster_set_params(sf_cluster* c, const sf_cluster_params* params) {
    return sf::guard([&] {
        SF_REQUIRE(c != nullptr && params != nullptr, SF_ERR_INVALID, "NULL argument");
        CL(c)->p = *params;
    });
}

extern "C" int sf_cluster_step(sf_cluster* c, float dt, int32_t n_steps) {
    return sf::guard([&] {
        SF_REQUIRE(c != nullptr && n_steps >= 0, SF_ERR_INVALID, "bad arguments");
        SF_REQUIRE(std::isfinite(dt), SF_ERR_INVALID, "dt must be finite");
        CL(c)->step_many(n_steps, dt);
    });
}

extern "C" int sf_cluster_synchronize(sf_cluster* c) {
    return sf::guard([&] {
        SF_REQUIRE(c != nullptr, SF_ERR_INVALID, "NULL argument");
        SF_HIP(hipStreamSynchronize(CL(c)->stream));
    });
}

extern "C" int sf_cluster_kernel_timing(sf_cluster* c, int32_t enable, double* avg_ms, int64_t* launches) {
    return sf::guard([&] {
        SF_REQUIRE(c != nullptr, SF_ERR_INVALID, "NULL argument");
        Cluster* k = CL(c);
        sf::KernelTimer& t = k->pair_timer;
        t.drain(k->stream);
        if (avg_ms) *avg_ms = t.launches ? t.ms / static_cast<double>(t.launches) : 0.0;
        if (launches) *launches = t.launches;
        t.ms = 0.0;
        t.launches = 0;
        t.enabled = enable != 0;
    });
}

extern "C" int sf_cluster_compute_forces(sf_cluster* c, float* forces) {
    return sf::guard([&] {
        SF_REQUIRE(c != nullptr && forces != nullptr, SF_ERR_INVALID, "NULL argument");
        Cluster* k = CL(c);
        if (k->n == 0) return;
        float* d = sf::dnew<float>(3 * static_cast<size_t>(k->n));
        try {
            k->forces_and_integrate(0.f, d, false);
            SF_HIP(hipMemcpyAsync(forces, d, 3 * k->n * sizeof(float), hipMemcpyDeviceToHost, k->stream));
            SF_HIP(hipStreamSynchronize(k->stream));
        } catch (...) {
            sf::dfree(d);
            throw;
        }
        sf::dfree(d);
    });
}

extern "C" int sf_cluster_num_bubbles(sf_cluster* c, int32_t* n) {
    return sf::guard([&] {
        SF_REQUIRE(c != nullptr && n != nullptr, SF_ERR_INVALID, "NULL argument");
        *n = CL(c)->n;
    });
}

extern "C" int sf_cluster_get_bubbles(sf_cluster* c, sf_cluster_bubble* out) {
    return sf::guard([&] {
        SF_REQUIRE(c != nullptr && out != nullptr, SF_ERR_INVALID, "NULL argument");
        Cluster* k = CL(c);
        std::vector<float4> hp, hv;
        k->download(hp, hv);
        for (int b = 0; b < k->n; ++b) {
            sf_cluster_bubble& o = out[b];
            o.id = k->ids[b];
            o.position[0] = hp[b].x; o.position[1] = hp[b].y; o.position[2] = hp[b].z;
            o.radius = hp[b].w;
            o.velocity[0] = hv[b].x; o.velocity[1] = hv[b].y; o.velocity[2] = hv[b].z;
            o.thickness_nm = hv[b].w;
            o.aspect_ratio = k->aspect[b];
            o.refractive_index = k->refr[b];
        }
    });
}

extern "C" int sf_cluster_set_bubbles(sf_cluster* c, int32_t n, const sf_cluster_bubble* bubbles) {
    return sf::guard([&] {
        SF_REQUIRE(c != nullptr, SF_ERR_INVALID, "NULL argument");
        check_bubbles(n, bubbles);
        Cluster* k = CL(c);
        std::vector<float4> hp(n), hv(n);
        k->ids.clear(); k->aspect.clear(); k->refr.clear();
        for (int b = 0; b < n; ++b) {
            hp[b] = make_float4(bubbles[b].position[0], bubbles[b].position[1], bubbles[b].position[2],
                                bubbles[b].radius);
            hv[b] = make_float4(bubbles[b].velocity[0], bubbles[b].velocity[1], bubbles[b].velocity[2],
                                bubbles[b].thickness_nm);
            k->ids.push_back(bubbles[b].id);
            k->aspect.push_back(bubbles[b].aspect_ratio);
            k->refr.push_back(bubbles[b].refractive_index);
            k->next_id = std::max(k->next_id, bubbles[b].id + 1);
        }
        k->upload(hp, hv);
        k->conn_valid = false;
    });
}

extern "C" int sf_cluster_num_connections(sf_cluster* c, int32_t* n) {
    return sf::guard([&] {
        SF_REQUIRE(c != nullptr && n != nullptr, SF_ERR_INVALID, "NULL argument");
        Cluster* k = CL(c);
        k->ensure_connections();
        *n = static_cast<int32_t>(k->conns.size());
    });
}

extern "C" int sf_cluster_get_connections(sf_cluster* c, sf_cluster_connection* out) {
    return sf::guard([&] {
        SF_REQUIRE(c != nullptr && out != nullptr, SF_ERR_INVALID, "NULL argument");
        Cluster* k = CL(c);
        k->ensure_connections();
        if (!k->conns.empty()) std::memcpy(out, k->conns.data(), k->conns.size() * sizeof(sf_cluster_connection));
    });
}

extern "C" int sf_cluster_neighbors(sf_cluster* c, int32_t* offsets, uint32_t* ids) {
    return sf::guard([&] {
        SF_REQUIRE(c != nullptr && offsets != nullptr, SF_ERR_INVALID, "NULL argument");
        Cluster* k = CL(c);
        k->ensure_connections();
        int32_t o = 0;
        for (int b = 0; b < k->n; ++b) {
            offsets[b] = o;
            const auto& nb = k->neighbors[b];
            if (ids && !nb.empty()) std::memcpy(ids + o, nb.data(), nb.size() * sizeof(uint32_t));
            o += static_cast<int32_t>(nb.size());
        }
        offsets[k->n] = o;
    });
}

extern "C" int sf_cluster_add_bubble(sf_cluster* c, const float position[3], float radius, uint32_t* id) {
    return sf::guard([&] {
        SF_REQUIRE(c != nullptr && position != nullptr, SF_ERR_INVALID, "NULL argument");
        SF_REQUIRE(radius > 0.f && std::isfinite(radius), SF_ERR_INVALID, "radius must be > 0");
        Cluster* k = CL(c);
        k->ensure_connections();   // the new bubble joins no connection until the next update
        std::vector<float4> hp, hv;
        k->download(hp, hv);
        const uint32_t nid = k->add_host(position, radius, hp, hv);
        k->upload(hp, hv);
        if (id) *id = nid;
    });
}

extern "C" int sf_cluster_remove_bubble(sf_cluster* c, uint32_t id, int32_t* removed) {
    return sf::guard([&] {
        SF_REQUIRE(c != nullptr, SF_ERR_INVALID, "NULL argument");
        Cluster* k = CL(c);
        k->ensure_connections();
        std::vector<float4> hp, hv;
        k->download(hp, hv);
        const bool ok = k->remove_host(id, hp, hv);
        if (ok) k->upload(hp, hv);
        if (removed) *removed = ok ? 1 : 0;
    });
}

extern "C" int sf_cluster_merge(sf_cluster* c, uint32_t id_a, uint32_t id_b, uint32_t* new_id) {
    return sf::guard([&] {
        SF_REQUIRE(c != nullptr, SF_ERR_INVALID, "NULL argument");
        CL(c)->ensure_connections();
        const uint32_t id = CL(c)->merge(id_a, id_b);
        if (new_id) *new_id = id;
    });
}

extern "C" int sf_cluster_update_connections(sf_cluster* c) {
    return sf::guard([&] {
        SF_REQUIRE(c != nullptr, SF_ERR_INVALID, "NULL argument");
        Cluster* k = CL(c);
        k->build_connections(k->contacts(false, nullptr));
    });
}

extern "C" int sf_cluster_foam_bubbles(sf_cluster* c, int32_t n_theta_min, int32_t n_theta_max, float radius_min,
                                       float radius_max, sf_bubble* out) {
    return sf::guard([&] {
        SF_REQUIRE(c != nullptr && out != nullptr, SF_ERR_INVALID, "NULL argument");
        SF_REQUIRE(n_theta_min >= 3 && n_theta_max >= n_theta_min, SF_ERR_INVALID, "bad n_theta range");
        SF_REQUIRE(radius_min > 0.f && radius_max > radius_min, SF_ERR_INVALID, "bad radius range");
        Cluster* k = CL(c);
        std::vector<float4> hp, hv;
        k->download(hp, hv);
        const double l0 = std::log(static_cast<double>(radius_min)), l1 = std::log(static_cast<double>(radius_max));
        for (int b = 0; b < k->n; ++b) {
            const double r = hp[b].w;
            double u = (std::log(r) - l0) / (l1 - l0);
            u = std::min(1.0, std::max(0.0, u));
            const int nt = static_cast<int>(std::lround(n_theta_min + u * (n_theta_max - n_theta_min)));
            out[b].n_theta = nt;
            out[b].n_phi = 2 * nt;
            out[b].bubble_radius = r;
            out[b].viscosity = k->p.viscosity;
            out[b].h0 = static_cast<double>(hv[b].w) * 1e-9;
            out[b].dt = 0.0;
        }
    });
}

extern "C" int sf_cluster_thickness_from_foam(sf_cluster* c, sf_foam* foam) {
    return sf::guard([&] {
        SF_REQUIRE(c != nullptr && foam != nullptr, SF_ERR_INVALID, "NULL argument");
        Cluster* k = CL(c);
        int32_t nf = 0;
        if (sf_foam_num_bubbles(foam, &nf) != SF_OK) throw sf::Error{SF_ERR_INVALID};
        SF_REQUIRE(nf == k->n, SF_ERR_INVALID, "foam has %d bubbles, cluster %d", nf, k->n);
        std::vector<double> mass(nf);
        std::vector<sf_bubble> fb(nf);
        if (sf_foam_stats(foam, mass.data(), nullptr, nullptr, nullptr) != SF_OK) throw sf::Error{SF_ERR_DEVICE};
        if (sf_foam_bubbles(foam, fb.data()) != SF_OK) throw sf::Error{SF_ERR_DEVICE};
        std::vector<float4> hp, hv;
        k->download(hp, hv);
        for (int b = 0; b < nf; ++b) {
            const double R = fb[b].bubble_radius;
            hv[b].w = static_cast<float>(mass[b] / (4.0 * 3.14159265358979323846 * R * R) * 1e9);
        }
        k->upload(hp, hv);
    });
}
