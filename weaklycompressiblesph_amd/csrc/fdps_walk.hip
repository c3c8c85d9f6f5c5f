// FDPS multiwalk drop-in: wcsph_hydro_dispatch_* / wcsph_hydro_retrieve_*.
//
// Replaces GPU_WC::HydrDispatchKernel / HydrRetrieveKernel
// (src/kernel.cu:289-409). Differences by design:
//  * records are packed in parallel into pinned SoA blocks, one block per
//    chunk of about 1/8 of the batch; each chunk's upload and kernel are
//    queued on the tag's stream as soon as it is packed, so packing the
//    next chunk overlaps the transfer and the force of this one (the
//    reference packs serially into AoS half records and issues four
//    synchronous copies, src/kernel.cu:317-373);
//  * positions and velocities are shipped in FP32 relative to the walk's
//    first i particle, the subtraction done in FP64 on the host (the
//    reference shifts by the walk's centre of mass and mean velocity,
//    src/kernel.cu:318-327, and stores FP16); pair differences are FP32;
//  * one workgroup per walk stages j tiles in LDS and each lane owns one i;
//    per-particle h and m are honoured (the reference assumes the first
//    particle's h and m for every pair, src/kernel.cu:303-304,400-401);
//  * the result copy is queued behind the last chunk's kernel, so retrieve
//    only waits and unpacks.
#include <chrono>
#include <mutex>

#include "common.hpp"
#include "pass.hpp"
#include "pair.hpp"

namespace wcsph {

namespace {

constexpr int kWalkThreads = 128;
constexpr int kWalkTile = 128;
constexpr int kMaxTags = 8;

template <int D>
struct WalkArgs {
    const int32_t* ioff;  // [n_walk+1], chunk-local
    const int32_t* joff;  // [n_walk+1], chunk-local
    const float* ix;      // [D][ni]: x - x0 of the walk
    const float* iv;      // [D][ni]: v - v0 of the walk
    const float* iq;      // [5][ni]: h rho pr c nu
    const float* jx;      // [D][nj]
    const float* jv;      // [D][nj]
    const float* jq;      // [6][nj]: h rho pr c nu m
    float* res;           // [ni_batch][D+2]: acc[D] dt div_v, at the batch's i index
    int64_t ni, nj;       // records of the chunk
    int64_t i_base;       // batch index of the chunk's first i
    PairConsts<float> k;
    float c_cfl;
};

template <int D>
__global__ __launch_bounds__(kWalkThreads) void k_walk_force(WalkArgs<D> a) {
    __shared__ float sx[D][kWalkTile];
    __shared__ float sv[D][kWalkTile];
    __shared__ float sq[6][kWalkTile];
    const int w = blockIdx.x;
    const int i0 = a.ioff[w], i1 = a.ioff[w + 1];
    const int j0 = a.joff[w], j1 = a.joff[w + 1];
    for (int ib = i0; ib < i1; ib += kWalkThreads) {
        const int i = ib + (int)threadIdx.x;
        const bool live = i < i1;
        float xi[D], vi[D];
        IState<float> I{1.f, 1.f, 0.f, 1.f, 0.f};
#pragma unroll
        for (int d = 0; d < D; ++d) {
            xi[d] = live ? a.ix[d * a.ni + i] : 0.f;
            vi[d] = live ? a.iv[d * a.ni + i] : 0.f;
        }
        if (live) {
            I.h = a.iq[0 * a.ni + i];
            I.rho = a.iq[1 * a.ni + i];
            I.pr = a.iq[2 * a.ni + i];
            I.c = a.iq[3 * a.ni + i];
            I.nu = a.iq[4 * a.ni + i];
        }
        float acc[D];
#pragma unroll
        for (int d = 0; d < D; ++d) acc[d] = 0.f;
        float divv = 0.f, vsig = 0.f;
        for (int jb = j0; jb < j1; jb += kWalkTile) {
            const int nt = min(kWalkTile, j1 - jb);
            __syncthreads();
            for (int t = threadIdx.x; t < nt; t += kWalkThreads) {
                const int j = jb + t;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    sx[d][t] = a.jx[d * a.nj + j];
                    sv[d][t] = a.jv[d * a.nj + j];
                }
#pragma unroll
                for (int q = 0; q < 6; ++q) sq[q][t] = a.jq[q * a.nj + j];
            }
            __syncthreads();
            if (!live) continue;
            for (int t = 0; t < nt; ++t) {
                float dr[D], dv[D];
                float r2 = 0.f, drdv = 0.f;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    dr[d] = xi[d] - sx[d][t];
                    dv[d] = vi[d] - sv[d][t];
                    r2 += dr[d] * dr[d];
                    drdv += dr[d] * dv[d];
                }
                if (r2 <= 0.f) continue;  // src/force.h:42
                const float hj = sq[0][t];
                float mu;
                const float den = pair_mu(r2, drdv, I, hj, a.k, mu, vsig);
                const float H = I.h + hj;
                if (r2 < H * H) {
                    JState<float> J{hj, sq[1][t], sq[2][t], sq[3][t], sq[4][t], sq[5][t]};
                    pair_force<D>(dr, dv, r2, drdv, den, mu, I, J, a.k, acc, divv);
                }
            }
        }
        if (live) {
            float* r = a.res + (a.i_base + i) * (D + 2);
#pragma unroll
            for (int d = 0; d < D; ++d) r[d] = acc[d];
            r[D] = a.c_cfl * I.h / (I.c + vsig);  // src/force.h:56
            r[D + 1] = divv / I.rho;              // src/force.h:57
        }
    }
}

constexpr int kMaxChunks = 8;

struct TagState {
    bool ready = false;
    bool outstanding = false;
    int dim = 0;
    hipStream_t stream = nullptr;       // kernels and the result copy
    hipStream_t copy_stream = nullptr;  // uploads, so chunk c + 1 moves while chunk c computes
    hipEvent_t ev_begin = nullptr, ev_end = nullptr;       // first upload .. result copy done
    hipEvent_t ev_k[2 * kMaxChunks] = {};                   // around each chunk's kernel
    hipEvent_t ev_up[kMaxChunks] = {};                      // each chunk's upload done
    int nchunks = 0;
    char* h_in = nullptr;
    size_t h_in_cap = 0;
    char* d_in = nullptr;
    size_t d_in_cap = 0;
    float* h_res = nullptr;
    float* d_res = nullptr;
    size_t res_cap = 0;
    std::vector<int32_t> ioff;  // offsets of the outstanding batch
    int64_t ni = 0;
    double pack_ms = 0.0, unpack_ms = 0.0;
};

struct Backend {
    std::mutex mu;
    bool inited = false;
    wcsph_params p{};
    TagState tags[kMaxTags];
};

Backend& backend() {
    static Backend b;
    return b;
}

void release_tag(TagState& t) {
    if (!t.ready) return;
    (void)hipStreamSynchronize(t.stream);
    if (t.h_in) (void)hipHostFree(t.h_in);
    if (t.d_in) (void)hipFree(t.d_in);
    if (t.h_res) (void)hipHostFree(t.h_res);
    if (t.d_res) (void)hipFree(t.d_res);
    (void)hipStreamSynchronize(t.copy_stream);
    (void)hipEventDestroy(t.ev_begin);
    (void)hipEventDestroy(t.ev_end);
    for (auto& e : t.ev_k) (void)hipEventDestroy(e);
    for (auto& e : t.ev_up) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(t.stream);
    (void)hipStreamDestroy(t.copy_stream);
    t = TagState{};
}

size_t align_up(size_t v, size_t a = 256) { return (v + a - 1) / a * a; }

// Byte layout of one staged chunk (identical on host and device).
struct Layout {
    size_t ioff, joff, ix, iv, iq, jx, jv, jq, total;
};

Layout layout(int D, int n_walk, int64_t ni, int64_t nj) {
    Layout L{};
    size_t o = 0;
    L.ioff = o; o = align_up(o + sizeof(int32_t) * (n_walk + 1));
    L.joff = o; o = align_up(o + sizeof(int32_t) * (n_walk + 1));
    L.ix = o; o = align_up(o + sizeof(float) * D * ni);
    L.iv = o; o = align_up(o + sizeof(float) * D * ni);
    L.iq = o; o = align_up(o + sizeof(float) * 5 * ni);
    L.jx = o; o = align_up(o + sizeof(float) * D * nj);
    L.jv = o; o = align_up(o + sizeof(float) * D * nj);
    L.jq = o; o = align_up(o + sizeof(float) * 6 * nj);
    L.total = o;
    return L;
}

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Pack walks [w0, w1) of the batch into one chunk block at h (SoA, FP32,
// positions and velocities relative to each walk's first i particle).
template <int D, class EPI, class EPJ>
void pack_chunk(char* h, const Layout& L, int w0, int w1, const std::vector<int32_t>& ioff,
                const std::vector<int32_t>& joff, const EPI* const* epi, const EPJ* const* epj) {
    const int nw = w1 - w0;
    const int64_t ni = ioff[w1] - ioff[w0], nj = joff[w1] - joff[w0];
    int32_t* lio = (int32_t*)(h + L.ioff);
    int32_t* ljo = (int32_t*)(h + L.joff);
    for (int w = 0; w <= nw; ++w) {
        lio[w] = ioff[w0 + w] - ioff[w0];
        ljo[w] = joff[w0 + w] - joff[w0];
    }
    float* ix = (float*)(h + L.ix);
    float* iv = (float*)(h + L.iv);
    float* iq = (float*)(h + L.iq);
    float* jx = (float*)(h + L.jx);
    float* jv = (float*)(h + L.jv);
    float* jq = (float*)(h + L.jq);
    parallel_for(nw, [&](int64_t k) {
        const int w = w0 + (int)k;
        const int n_i = ioff[w + 1] - ioff[w], n_j = joff[w + 1] - joff[w];
        double x0[D], v0[D];
        for (int d = 0; d < D; ++d) {
            x0[d] = n_i ? epi[w][0].pos[d] : 0.0;
            v0[d] = n_i ? epi[w][0].vel[d] : 0.0;
        }
        for (int t = 0; t < n_i; ++t) {
            const EPI& e = epi[w][t];
            const int64_t i = ioff[w] - ioff[w0] + t;
            for (int d = 0; d < D; ++d) {
                ix[d * ni + i] = (float)(e.pos[d] - x0[d]);
                iv[d * ni + i] = (float)(e.vel[d] - v0[d]);
            }
            iq[0 * ni + i] = (float)e.smth;
            iq[1 * ni + i] = (float)e.dens;
            iq[2 * ni + i] = (float)(e.pres / (e.dens * e.dens));
            iq[3 * ni + i] = (float)e.snds;
            iq[4 * ni + i] = (float)e.visc;
        }
        for (int t = 0; t < n_j; ++t) {
            const EPJ& e = epj[w][t];
            const int64_t j = joff[w] - joff[w0] + t;
            for (int d = 0; d < D; ++d) {
                jx[d * nj + j] = (float)(e.pos[d] - x0[d]);
                jv[d * nj + j] = (float)(e.vel[d] - v0[d]);
            }
            jq[0 * nj + j] = (float)e.smth;
            jq[1 * nj + j] = (float)e.dens;
            jq[2 * nj + j] = (float)(e.pres / (e.dens * e.dens));
            jq[3 * nj + j] = (float)e.snds;
            jq[4 * nj + j] = (float)e.visc;
            jq[5 * nj + j] = (float)e.mass;
        }
    }, 0, 64);  // walks are heavy: spread from 64 walks up
}

template <int D, class EPI, class EPJ>
void dispatch(int32_t tag, int32_t n_walk, const EPI* const* epi, const int32_t* n_epi,
              const EPJ* const* epj, const int32_t* n_epj) {
    const auto t_start = std::chrono::steady_clock::now();
    Backend& B = backend();
    std::lock_guard<std::mutex> lk(B.mu);
    WCSPH_REQUIRE(B.inited, WCSPH_ESTATE, "wcsph_init has not been called");
    WCSPH_REQUIRE(B.p.dim == D, WCSPH_EINVAL, "dispatch dimension differs from wcsph_init's dim");
    WCSPH_REQUIRE(tag >= 0 && tag < kMaxTags, WCSPH_EINVAL, "tag out of range [0, 8)");
    WCSPH_REQUIRE(n_walk >= 0, WCSPH_EINVAL, "n_walk < 0");
    WCSPH_REQUIRE(n_walk == 0 || (epi && n_epi && epj && n_epj), WCSPH_EINVAL, "null walk arrays");
    TagState& T = B.tags[tag];
    WCSPH_REQUIRE(!T.outstanding, WCSPH_ESTATE, "previous batch of this tag was not retrieved");
    if (!T.ready) {
        WCSPH_HIP(hipSetDevice(B.p.device));
        WCSPH_HIP(hipStreamCreateWithFlags(&T.stream, hipStreamNonBlocking));
        WCSPH_HIP(hipStreamCreateWithFlags(&T.copy_stream, hipStreamNonBlocking));
        WCSPH_HIP(hipEventCreate(&T.ev_begin));
        WCSPH_HIP(hipEventCreate(&T.ev_end));
        for (auto& e : T.ev_k) WCSPH_HIP(hipEventCreate(&e));
        for (auto& e : T.ev_up) WCSPH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        T.ready = true;
    }
    WCSPH_HIP(hipSetDevice(B.p.device));
    // uploads of a dispatch that failed part way may still read the staging
    // buffer: let them finish before it is packed again
    WCSPH_HIP(hipStreamSynchronize(T.copy_stream));
    std::vector<int32_t> ioff(n_walk + 1), joff(n_walk + 1);
    int64_t ni = 0, nj = 0;
    ioff[0] = joff[0] = 0;
    for (int w = 0; w < n_walk; ++w) {
        WCSPH_REQUIRE(n_epi[w] >= 0 && n_epj[w] >= 0, WCSPH_EINVAL, "negative walk size");
        WCSPH_REQUIRE(n_epi[w] == 0 || epi[w], WCSPH_EINVAL, "null epi walk pointer");
        WCSPH_REQUIRE(n_epj[w] == 0 || epj[w], WCSPH_EINVAL, "null epj walk pointer");
        ni += n_epi[w];
        nj += n_epj[w];
        WCSPH_REQUIRE(ni < INT32_MAX && nj < INT32_MAX, WCSPH_EOVERFLOW, "batch exceeds 2^31 records");
        ioff[w + 1] = (int32_t)ni;
        joff[w + 1] = (int32_t)nj;
    }
    // chunks of about equal record counts (at least 256 walks each)
    const int nch = std::max(1, std::min(kMaxChunks, n_walk / 256));
    std::vector<int> wb(nch + 1, n_walk);
    wb[0] = 0;
    {
        const int64_t tot = ni + nj;
        int c = 1;
        for (int w = 0; w < n_walk && c < nch; ++w)
            if (((int64_t)ioff[w] + joff[w]) * nch >= tot * c) wb[c++] = w;
        for (; c < nch; ++c) wb[c] = n_walk;
    }
    std::vector<Layout> Ls(nch);
    std::vector<size_t> base(nch + 1, 0);
    for (int c = 0; c < nch; ++c) {
        Ls[c] = layout(D, wb[c + 1] - wb[c], ioff[wb[c + 1]] - ioff[wb[c]], joff[wb[c + 1]] - joff[wb[c]]);
        base[c + 1] = base[c] + Ls[c].total;
    }
    const size_t total = std::max<size_t>(base[nch], 256);
    if (total > T.h_in_cap) {
        if (T.h_in) WCSPH_HIP(hipHostFree(T.h_in));
        if (T.d_in) WCSPH_HIP(hipFree(T.d_in));
        T.h_in = nullptr;
        T.d_in = nullptr;
        const size_t cap = total + total / 4;
        WCSPH_HIP(hipHostMalloc((void**)&T.h_in, cap, hipHostMallocDefault));
        WCSPH_HIP(hipMalloc((void**)&T.d_in, cap));
        T.h_in_cap = T.d_in_cap = cap;
    }
    const size_t res_bytes = sizeof(float) * (D + 2) * std::max<int64_t>(ni, 1);
    if (res_bytes > T.res_cap) {
        if (T.h_res) WCSPH_HIP(hipHostFree(T.h_res));
        if (T.d_res) WCSPH_HIP(hipFree(T.d_res));
        T.h_res = nullptr;
        T.d_res = nullptr;
        const size_t cap = res_bytes + res_bytes / 4;
        WCSPH_HIP(hipHostMalloc((void**)&T.h_res, cap, hipHostMallocDefault));
        WCSPH_HIP(hipMalloc((void**)&T.d_res, cap));
        T.res_cap = cap;
    }
    WalkArgs<D> a;
    a.res = T.d_res;
    a.k.alpha = (float)B.p.alpha;
    a.k.beta = (float)B.p.beta;
    a.k.eps = (float)B.p.eps;
    a.k.cnorm = (float)(D == 2 ? 80.0 / (7.0 * M_PI) : 16.0 / M_PI);
    a.k.vsig_mode = B.p.vsig_mode;
    a.c_cfl = (float)B.p.c_cfl;
    WCSPH_HIP(hipEventRecord(T.ev_begin, T.copy_stream));
    double pack_ms = 0.0;
    for (int c = 0; c < nch; ++c) {
        const int w0 = wb[c], w1 = wb[c + 1];
        const Layout& L = Ls[c];
        char* h = T.h_in + base[c];
        const auto t0 = std::chrono::steady_clock::now();
        pack_chunk<D>(h, L, w0, w1, ioff, joff, epi, epj);
        pack_ms += ms_since(t0);
        // the chunk goes to the device while the next one is packed, and its
        // kernel waits only for its own upload
        WCSPH_HIP(hipMemcpyAsync(T.d_in + base[c], h, L.total, hipMemcpyHostToDevice, T.copy_stream));
        WCSPH_HIP(hipEventRecord(T.ev_up[c], T.copy_stream));
        WCSPH_HIP(hipStreamWaitEvent(T.stream, T.ev_up[c], 0));
        char* d = T.d_in + base[c];
        a.ioff = (const int32_t*)(d + L.ioff);
        a.joff = (const int32_t*)(d + L.joff);
        a.ix = (const float*)(d + L.ix);
        a.iv = (const float*)(d + L.iv);
        a.iq = (const float*)(d + L.iq);
        a.jx = (const float*)(d + L.jx);
        a.jv = (const float*)(d + L.jv);
        a.jq = (const float*)(d + L.jq);
        a.ni = ioff[w1] - ioff[w0];
        a.nj = joff[w1] - joff[w0];
        a.i_base = ioff[w0];
        WCSPH_HIP(hipEventRecord(T.ev_k[2 * c], T.stream));
        if (w1 > w0) hipLaunchKernelGGL(k_walk_force<D>, dim3(w1 - w0), dim3(kWalkThreads), 0, T.stream, a);
        WCSPH_HIP(hipGetLastError());
        WCSPH_HIP(hipEventRecord(T.ev_k[2 * c + 1], T.stream));
    }
    if (ni > 0)
        WCSPH_HIP(hipMemcpyAsync(T.h_res, T.d_res, sizeof(float) * (D + 2) * ni, hipMemcpyDeviceToHost,
                                 T.stream));
    WCSPH_HIP(hipEventRecord(T.ev_end, T.stream));
    T.nchunks = nch;
    T.ioff.swap(ioff);
    T.ni = ni;
    T.dim = D;
    T.pack_ms = pack_ms;
    (void)t_start;
    T.outstanding = true;
}

template <int D, class RES>
void retrieve(int32_t tag, int32_t n_walk, const int32_t* ni_w, RES** force) {
    Backend& B = backend();
    std::lock_guard<std::mutex> lk(B.mu);
    WCSPH_REQUIRE(B.inited, WCSPH_ESTATE, "wcsph_init has not been called");
    WCSPH_REQUIRE(tag >= 0 && tag < kMaxTags, WCSPH_EINVAL, "tag out of range [0, 8)");
    TagState& T = B.tags[tag];
    WCSPH_REQUIRE(T.outstanding, WCSPH_ESTATE, "no dispatched batch on this tag");
    WCSPH_REQUIRE(T.dim == D, WCSPH_EINVAL, "retrieve dimension differs from dispatch");
    WCSPH_REQUIRE(n_walk + 1 == (int32_t)T.ioff.size(), WCSPH_EINVAL, "n_walk differs from dispatch");
    for (int w = 0; w < n_walk; ++w)
        WCSPH_REQUIRE(ni_w[w] == T.ioff[w + 1] - T.ioff[w], WCSPH_EINVAL, "ni differs from dispatch");
    WCSPH_REQUIRE(n_walk == 0 || force, WCSPH_EINVAL, "null force array");
    T.outstanding = false;  // the batch is consumed whatever happens below
    WCSPH_HIP(hipStreamSynchronize(T.stream));
    const auto t0 = std::chrono::steady_clock::now();
    const float* r = T.h_res;
    const std::vector<int32_t>& off = T.ioff;
    parallel_for(n_walk, [&](int64_t w) {
        RES* f = force[w];
        for (int t = 0; t < ni_w[w]; ++t) {
            const float* q = r + (int64_t)(off[w] + t) * (D + 2);
            for (int d = 0; d < D; ++d) f[t].acc[d] = q[d];
            f[t].dt = q[D];
            f[t].div_v = q[D + 1];
        }
    }, 0, 64);
    T.unpack_ms = ms_since(t0);
}

int backend_last_timing(int32_t tag, double* out) {
    return guarded([&] {
        WCSPH_REQUIRE(tag >= 0 && tag < kMaxTags && out, WCSPH_EINVAL, "bad tag or null output");
        Backend& B = backend();
        std::lock_guard<std::mutex> lk(B.mu);
        TagState& T = B.tags[tag];
        WCSPH_REQUIRE(T.ready && T.nchunks > 0, WCSPH_ESTATE, "tag has no batch yet");
        WCSPH_HIP(hipEventSynchronize(T.ev_end));
        float ms = 0.f, k = 0.f;
        for (int c = 0; c < T.nchunks; ++c) {
            WCSPH_HIP(hipEventElapsedTime(&ms, T.ev_k[2 * c], T.ev_k[2 * c + 1]));
            k += ms;
        }
        WCSPH_HIP(hipEventElapsedTime(&ms, T.ev_begin, T.ev_end));
        out[0] = T.pack_ms;
        out[1] = ms;
        out[2] = k;
        out[3] = T.unpack_ms;
        out[4] = T.nchunks;
    });
}
}  // namespace

int backend_init(const wcsph_params* p) {
    return guarded([&] {
        WCSPH_REQUIRE(p != nullptr, WCSPH_EINVAL, "null params");
        WCSPH_REQUIRE(p->dim == 2 || p->dim == 3, WCSPH_EINVAL, "dim must be 2 or 3");
        Backend& B = backend();
        std::lock_guard<std::mutex> lk(B.mu);
        int ndev = 0;
        WCSPH_HIP(hipGetDeviceCount(&ndev));
        WCSPH_REQUIRE(p->device >= 0 && p->device < ndev, WCSPH_EINVAL, "device ordinal out of range");
        for (auto& t : B.tags) WCSPH_REQUIRE(!t.outstanding, WCSPH_ESTATE, "a batch is still outstanding");
        if (B.inited && B.p.device != p->device)
            for (auto& t : B.tags) release_tag(t);
        B.p = *p;
        B.inited = true;
        // set-up of the no-allocation force pass on this device
        (void)noalloc_ptr(p->device);
    });
}

int backend_finalize() {
    return guarded([&] {
        Backend& B = backend();
        std::lock_guard<std::mutex> lk(B.mu);
        for (auto& t : B.tags) release_tag(t);
        B.inited = false;
    });
}

int backend_last_ms(int32_t tag, float* ms) {
    double t[5];
    const int rc = backend_last_timing(tag, t);
    if (rc == WCSPH_OK) *ms = (float)t[2];
    return rc;
}

}  // namespace wcsph

using namespace wcsph;

extern "C" {

int wcsph_init(const wcsph_params* p) { return backend_init(p); }
int wcsph_finalize(void) { return backend_finalize(); }
int wcsph_hydro_last_kernel_ms(int32_t tag, float* ms) {
    if (!ms) {
        set_error("null output");
        return WCSPH_EINVAL;
    }
    return backend_last_ms(tag, ms);
}
int wcsph_hydro_last_timing(int32_t tag, double* out) { return backend_last_timing(tag, out); }

int wcsph_hydro_dispatch_3d(int32_t tag, int32_t n_walk, const wcsph_epi_hydro_3d* const* epi,
                            const int32_t* n_epi, const wcsph_epj_hydro_3d* const* epj,
                            const int32_t* n_epj) {
    return guarded([&] { dispatch<3>(tag, n_walk, epi, n_epi, epj, n_epj); });
}
int wcsph_hydro_dispatch_2d(int32_t tag, int32_t n_walk, const wcsph_epi_hydro_2d* const* epi,
                            const int32_t* n_epi, const wcsph_epj_hydro_2d* const* epj,
                            const int32_t* n_epj) {
    return guarded([&] { dispatch<2>(tag, n_walk, epi, n_epi, epj, n_epj); });
}
int wcsph_hydro_retrieve_3d(int32_t tag, int32_t n_walk, const int32_t* ni, wcsph_result_hydro_3d** force) {
    return guarded([&] { retrieve<3>(tag, n_walk, ni, force); });
}
int wcsph_hydro_retrieve_2d(int32_t tag, int32_t n_walk, const int32_t* ni, wcsph_result_hydro_2d** force) {
    return guarded([&] { retrieve<2>(tag, n_walk, ni, force); });
}

}  // extern "C"
