// FDPS multiwalk drop-in: wcsph_hydro_dispatch_* / wcsph_hydro_retrieve_*.
//
// Replaces GPU_WC::HydrDispatchKernel / HydrRetrieveKernel
// (src/kernel.cu:289-409). Differences by design:
//  * records are packed in parallel into one pinned SoA block and uploaded
//    with a single async copy on the tag's stream (the reference packs
//    serially into AoS half records and issues four synchronous copies,
//    src/kernel.cu:317-373);
//  * positions and velocities stay FP64 on the device; pair differences are
//    formed in FP64 and rounded once to FP32, so no per-walk shift or
//    nondimensionalisation is needed (src/kernel.cu:318-368);
//  * one workgroup per walk stages j tiles in LDS and each lane owns one i;
//    per-particle h and m are honoured (the reference assumes the first
//    particle's h and m for every pair, src/kernel.cu:303-304,400-401);
//  * the result copy is queued behind the kernel, so retrieve only waits
//    and unpacks.
#include <mutex>

#include "common.hpp"
#include "pair.hpp"

namespace wcsph {

namespace {

constexpr int kWalkThreads = 128;
constexpr int kWalkTile = 128;
constexpr int kMaxTags = 8;

template <int D>
struct WalkArgs {
    const int32_t* ioff;  // [n_walk+1]
    const int32_t* joff;  // [n_walk+1]
    const double* ix;     // [D][ni]
    const double* iv;     // [D][ni]
    const float* iq;      // [5][ni]: h rho pr c nu
    const double* jx;     // [D][nj]
    const double* jv;     // [D][nj]
    const float* jq;      // [6][nj]: h rho pr c nu m
    float* res;           // [D+2][ni]: acc[D] dt div_v
    int64_t ni, nj;
    PairConsts<float> k;
    float c_cfl;
};

template <int D>
__global__ __launch_bounds__(kWalkThreads) void k_walk_force(WalkArgs<D> a) {
    __shared__ double sx[D][kWalkTile];
    __shared__ double sv[D][kWalkTile];
    __shared__ float sq[6][kWalkTile];
    const int w = blockIdx.x;
    const int i0 = a.ioff[w], i1 = a.ioff[w + 1];
    const int j0 = a.joff[w], j1 = a.joff[w + 1];
    for (int ib = i0; ib < i1; ib += kWalkThreads) {
        const int i = ib + (int)threadIdx.x;
        const bool live = i < i1;
        double xi[D], vi[D];
        IState<float> I{1.f, 1.f, 0.f, 1.f, 0.f};
#pragma unroll
        for (int d = 0; d < D; ++d) {
            xi[d] = live ? a.ix[d * a.ni + i] : 0.0;
            vi[d] = live ? a.iv[d * a.ni + i] : 0.0;
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
                    dr[d] = (float)(xi[d] - sx[d][t]);
                    dv[d] = (float)(vi[d] - sv[d][t]);
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
#pragma unroll
            for (int d = 0; d < D; ++d) a.res[d * a.ni + i] = acc[d];
            a.res[D * a.ni + i] = a.c_cfl * I.h / (I.c + vsig);  // src/force.h:56
            a.res[(D + 1) * a.ni + i] = divv / I.rho;           // src/force.h:57
        }
    }
}

struct TagState {
    bool ready = false;
    bool outstanding = false;
    int dim = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    char* h_in = nullptr;
    size_t h_in_cap = 0;
    char* d_in = nullptr;
    size_t d_in_cap = 0;
    float* h_res = nullptr;
    float* d_res = nullptr;
    size_t res_cap = 0;
    std::vector<int32_t> ioff;  // offsets of the outstanding batch
    int64_t ni = 0;
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
    (void)hipEventDestroy(t.ev0);
    (void)hipEventDestroy(t.ev1);
    (void)hipStreamDestroy(t.stream);
    t = TagState{};
}

size_t align_up(size_t v, size_t a = 256) { return (v + a - 1) / a * a; }

// Byte layout of one staged batch (identical on host and device).
struct Layout {
    size_t ioff, joff, ix, iv, iq, jx, jv, jq, total;
};

Layout layout(int D, int n_walk, int64_t ni, int64_t nj) {
    Layout L{};
    size_t o = 0;
    L.ioff = o; o = align_up(o + sizeof(int32_t) * (n_walk + 1));
    L.joff = o; o = align_up(o + sizeof(int32_t) * (n_walk + 1));
    L.ix = o; o = align_up(o + sizeof(double) * D * ni);
    L.iv = o; o = align_up(o + sizeof(double) * D * ni);
    L.iq = o; o = align_up(o + sizeof(float) * 5 * ni);
    L.jx = o; o = align_up(o + sizeof(double) * D * nj);
    L.jv = o; o = align_up(o + sizeof(double) * D * nj);
    L.jq = o; o = align_up(o + sizeof(float) * 6 * nj);
    L.total = o;
    return L;
}

template <int D, class EPI, class EPJ>
void dispatch(int32_t tag, int32_t n_walk, const EPI* const* epi, const int32_t* n_epi,
              const EPJ* const* epj, const int32_t* n_epj) {
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
        WCSPH_HIP(hipEventCreate(&T.ev0));
        WCSPH_HIP(hipEventCreate(&T.ev1));
        T.ready = true;
    }
    WCSPH_HIP(hipSetDevice(B.p.device));
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
    const Layout L = layout(D, n_walk, ni, nj);
    if (L.total > T.h_in_cap) {
        if (T.h_in) WCSPH_HIP(hipHostFree(T.h_in));
        if (T.d_in) WCSPH_HIP(hipFree(T.d_in));
        T.h_in = nullptr;
        T.d_in = nullptr;
        const size_t cap = L.total + L.total / 4;
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
    char* h = T.h_in;
    std::copy(ioff.begin(), ioff.end(), (int32_t*)(h + L.ioff));
    std::copy(joff.begin(), joff.end(), (int32_t*)(h + L.joff));
    double* ix = (double*)(h + L.ix);
    double* iv = (double*)(h + L.iv);
    float* iq = (float*)(h + L.iq);
    double* jx = (double*)(h + L.jx);
    double* jv = (double*)(h + L.jv);
    float* jq = (float*)(h + L.jq);
    parallel_for(n_walk, [&](int64_t w) {
        for (int t = 0; t < n_epi[w]; ++t) {
            const EPI& e = epi[w][t];
            const int64_t i = ioff[w] + t;
            for (int d = 0; d < D; ++d) {
                ix[d * ni + i] = e.pos[d];
                iv[d * ni + i] = e.vel[d];
            }
            iq[0 * ni + i] = (float)e.smth;
            iq[1 * ni + i] = (float)e.dens;
            iq[2 * ni + i] = (float)(e.pres / (e.dens * e.dens));
            iq[3 * ni + i] = (float)e.snds;
            iq[4 * ni + i] = (float)e.visc;
        }
        for (int t = 0; t < n_epj[w]; ++t) {
            const EPJ& e = epj[w][t];
            const int64_t j = joff[w] + t;
            for (int d = 0; d < D; ++d) {
                jx[d * nj + j] = e.pos[d];
                jv[d * nj + j] = e.vel[d];
            }
            jq[0 * nj + j] = (float)e.smth;
            jq[1 * nj + j] = (float)e.dens;
            jq[2 * nj + j] = (float)(e.pres / (e.dens * e.dens));
            jq[3 * nj + j] = (float)e.snds;
            jq[4 * nj + j] = (float)e.visc;
            jq[5 * nj + j] = (float)e.mass;
        }
    }, n_walk >= 64 ? 0 : 1);
    WCSPH_HIP(hipMemcpyAsync(T.d_in, T.h_in, L.total, hipMemcpyHostToDevice, T.stream));
    WalkArgs<D> a;
    char* d = T.d_in;
    a.ioff = (const int32_t*)(d + L.ioff);
    a.joff = (const int32_t*)(d + L.joff);
    a.ix = (const double*)(d + L.ix);
    a.iv = (const double*)(d + L.iv);
    a.iq = (const float*)(d + L.iq);
    a.jx = (const double*)(d + L.jx);
    a.jv = (const double*)(d + L.jv);
    a.jq = (const float*)(d + L.jq);
    a.res = T.d_res;
    a.ni = ni;
    a.nj = nj;
    a.k.alpha = (float)B.p.alpha;
    a.k.beta = (float)B.p.beta;
    a.k.eps = (float)B.p.eps;
    a.k.cnorm = (float)(D == 2 ? 80.0 / (7.0 * M_PI) : 16.0 / M_PI);
    a.k.vsig_mode = B.p.vsig_mode;
    a.c_cfl = (float)B.p.c_cfl;
    WCSPH_HIP(hipEventRecord(T.ev0, T.stream));
    if (n_walk > 0) {
        hipLaunchKernelGGL(k_walk_force<D>, dim3(n_walk), dim3(kWalkThreads), 0, T.stream, a);
        WCSPH_HIP(hipGetLastError());
    }
    WCSPH_HIP(hipEventRecord(T.ev1, T.stream));
    if (ni > 0)
        WCSPH_HIP(hipMemcpyAsync(T.h_res, T.d_res, sizeof(float) * (D + 2) * ni, hipMemcpyDeviceToHost,
                                 T.stream));
    T.ioff.swap(ioff);
    T.ni = ni;
    T.dim = D;
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
    const float* r = T.h_res;
    const int64_t ni = T.ni;
    const std::vector<int32_t>& off = T.ioff;
    parallel_for(n_walk, [&](int64_t w) {
        RES* f = force[w];
        for (int t = 0; t < ni_w[w]; ++t) {
            const int64_t i = off[w] + t;
            for (int d = 0; d < D; ++d) f[t].acc[d] = r[d * ni + i];
            f[t].dt = r[D * ni + i];
            f[t].div_v = r[(D + 1) * ni + i];
        }
    }, n_walk >= 64 ? 0 : 1);
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
    return guarded([&] {
        WCSPH_REQUIRE(tag >= 0 && tag < kMaxTags && ms, WCSPH_EINVAL, "bad tag or null output");
        Backend& B = backend();
        std::lock_guard<std::mutex> lk(B.mu);
        TagState& T = B.tags[tag];
        WCSPH_REQUIRE(T.ready, WCSPH_ESTATE, "tag has no batch yet");
        WCSPH_HIP(hipEventSynchronize(T.ev1));
        WCSPH_HIP(hipEventElapsedTime(ms, T.ev0, T.ev1));
    });
}

}  // namespace wcsph

using namespace wcsph;

extern "C" {

int wcsph_init(const wcsph_params* p) { return backend_init(p); }
int wcsph_finalize(void) { return backend_finalize(); }
int wcsph_hydro_last_kernel_ms(int32_t tag, float* ms) { return backend_last_ms(tag, ms); }

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
