// FDPS multiwalk drop-in: wcsph_hydro_dispatch_* / wcsph_hydro_retrieve_*.
//
// Replaces GPU_WC::HydrDispatchKernel / HydrRetrieveKernel
// (src/kernel.cu:289-409). Differences by design:
//  * records are packed in parallel into pinned SoA blocks, one block per
//    chunk of about 1/8 of the batch; each chunk's upload is queued on a
//    copy stream as soon as it is packed and its kernel on one of two
//    kernel streams (even and odd chunks alternate), so packing the next
//    chunk overlaps the transfer and the force of this one, and a chunk's
//    tail of long walks overlaps the next chunk's kernel (the reference
//    packs serially into AoS half records and issues four synchronous
//    copies, src/kernel.cu:317-373);
//  * positions and velocities are shipped in FP32 relative to the walk's
//    first i particle, the subtraction done in FP64 on the host (the
//    reference shifts by the walk's centre of mass and mean velocity,
//    src/kernel.cu:318-327, and stores FP16); pair differences are FP32;
//  * one workgroup per walk (long walks cut into pieces along the j list,
//    launched longest first) stages j tiles in LDS; each i is owned by
//    256 / (number of i in the walk) lanes that split its j list; each lane
//    lists its neighbours by the support test alone, then runs the pair
//    body over the list, and all partial sums are added in a fixed order
//    (the reference gives one thread per i and evaluates every candidate,
//    src/kernel.cu:241-280); per-particle h and m are honoured (the
//    reference assumes the first particle's h and m for every pair,
//    src/kernel.cu:303-304,400-401);
//  * the result copy is queued behind the last kernels of both streams, so
//    retrieve only waits and unpacks.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <utility>
#include <vector>

#include "common.hpp"
#include "pass.hpp"
#include "pair.hpp"

namespace wcsph {

namespace {

constexpr int kWalkThreads = 256;
constexpr int kWalkTile = 256;  // j records per LDS tile (list entries are one byte)
constexpr int kWalkCap = 64;    // neighbours listed per lane before the pair body runs
constexpr int kWalkScan = 4;    // candidates tested per scan step
constexpr int kWalkPiece = 128; // serial candidates per lane above which a walk is cut into pieces
constexpr int kWalkMinPieceJ = 64;  // fewest j candidates of a piece
constexpr int kMaxTags = 8;

template <int D>
struct WalkArgs {
    const int32_t* ioff;  // [n_walk+1], chunk-local
    const int32_t* joff;  // [n_walk+1], chunk-local
    // work items, longest first: walk, j range [jb, je) (chunk-local), and
    // where the sums go (-1: final results in res; else the float offset of
    // this piece's partial sums in part)
    const int32_t* it_walk;
    const int32_t* it_jb;
    const int32_t* it_je;
    const int32_t* it_out;
    // walks cut into pieces: walk, float offset of its partials, pieces
    const int32_t* sp_walk;
    const int32_t* sp_base;
    const int32_t* sp_np;
    float* part;          // partial sums of pieces: [piece][i of walk][D+2]
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

// One workgroup of kWalkThreads per work item: a walk, or one piece of a
// long walk's j list (its sums then go to a partial buffer that
// k_walk_combine adds up in piece order). A round takes nr <= kWalkThreads
// of the walk's i particles and splits each one's j list over
// S = kWalkThreads / nr threads ("parts"): thread t takes i = t % nr and
// every S-th candidate of each tile from t / nr on. The parts' sums are
// combined in part order through LDS at the end of the round, so the
// result does not depend on scheduling. j records are staged in LDS as
// float4 {x, y, z, h}, {vx, vy, vz, m}, {rho, P/rho^2, c, nu}.
//
// The superset lists hold about 7 candidates per neighbour, and the lanes of
// a wave (64 different i against the same j) seldom agree on which are
// neighbours, so each lane first scans its candidates with the support test
// alone and appends the tile index of every neighbour to a list of its own
// in LDS (kWalkCap entries, one byte each); the pair body then runs over the
// list, every lane on a neighbour of its own. Neighbours are taken in
// candidate order, so the sums are those of a plain loop. In v_sig mode 0
// (src/force.h:47, where v_sig stays 0) nothing outside the support is
// evaluated; in mode 1 (src/kernel.cu:263) -mu enters v_sig for every
// candidate during the scan, as the reference kernel takes it over the
// whole list.
template <int D, int VS>
__global__ __launch_bounds__(kWalkThreads) void k_walk_force(WalkArgs<D> a) {
    __shared__ float4 stage[3][kWalkTile];
    __shared__ uint8_t lst[kWalkCap][kWalkThreads];  // lane t's list is column t
    float4* sx = stage[0];
    float4* sv = stage[1];
    float4* sq = stage[2];
    // the parts' sums (acc[D], divv, vsig of each thread) reuse sv and sq
    static_assert((D + 2) * kWalkThreads * sizeof(float) <= 2 * kWalkTile * sizeof(float4), "red fits sv, sq");
    float(*red)[kWalkThreads] = reinterpret_cast<float(*)[kWalkThreads]>(&stage[1][0]);
    const int w = a.it_walk[blockIdx.x];
    const int i0 = a.ioff[w], i1 = a.ioff[w + 1];
    const int j0 = a.it_jb[blockIdx.x], j1 = a.it_je[blockIdx.x];
    const int out = a.it_out[blockIdx.x];
    const int t = (int)threadIdx.x;
    for (int ib = i0; ib < i1; ib += kWalkThreads) {
        const int nr = min(kWalkThreads, i1 - ib);  // i particles of this round (block-uniform)
        const int S = kWalkThreads / nr;             // parts per i
        const int part = t / nr;
        const int i = ib + t % nr;
        const bool live = part < S;
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
            for (int u = t; u < nt; u += kWalkThreads) {
                const int j = jb + u;
                float x[3] = {0.f, 0.f, 0.f}, v[3] = {0.f, 0.f, 0.f};
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    x[d] = a.jx[d * a.nj + j];
                    v[d] = a.jv[d * a.nj + j];
                }
                sx[u] = make_float4(x[0], x[1], x[2], a.jq[0 * a.nj + j]);
                sv[u] = make_float4(v[0], v[1], v[2], a.jq[5 * a.nj + j]);
                sq[u] = make_float4(a.jq[1 * a.nj + j], a.jq[2 * a.nj + j], a.jq[3 * a.nj + j], a.jq[4 * a.nj + j]);
            }
            __syncthreads();
            // this lane's candidates: every S-th of the tile from its part on
            int u = live ? part : nt;
            // the support test of candidate v (appends v to the list; r2 = 0,
            // the i particle itself, is listed and contributes nothing)
            auto test = [&](int v, const float4& pj, int& cnt) {
                float r2 = 0.f;
                float dr[D];
                const float pjv[3] = {pj.x, pj.y, pj.z};
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    dr[d] = xi[d] - pjv[d];
                    r2 += dr[d] * dr[d];
                }
                if (VS == 1 && r2 > 0.f) {
                    const float4 vj = sv[v];
                    const float vjv[3] = {vj.x, vj.y, vj.z};
                    float drdv = 0.f;
#pragma unroll
                    for (int d = 0; d < D; ++d) drdv += dr[d] * (vi[d] - vjv[d]);
                    float mu;
                    (void)pair_mu(r2, drdv, I, pj.w, a.k, mu, vsig);
                }
                lst[cnt][t] = (uint8_t)v;  // kept only if counted
                cnt += r2 < pair_support2(I.h, pj.w, a.k.grad_mode) ? 1 : 0;
            };
            while (u < nt) {
                // scan in groups of kWalkScan candidates (their LDS reads in
                // flight together) until the list is nearly full, then the
                // last few one by one
                int cnt = 0;
                for (; u + (kWalkScan - 1) * S < nt && cnt <= kWalkCap - kWalkScan; u += kWalkScan * S) {
                    float4 pj[kWalkScan];
#pragma unroll
                    for (int q = 0; q < kWalkScan; ++q) pj[q] = sx[u + q * S];
#pragma unroll
                    for (int q = 0; q < kWalkScan; ++q) test(u + q * S, pj[q], cnt);
                }
                if (u + (kWalkScan - 1) * S >= nt)
                    for (; u < nt && cnt < kWalkCap; u += S) test(u, sx[u], cnt);
                // pair body over the list
                for (int k = 0; k < cnt; ++k) {
                    const int v = lst[k][t];
                    const float4 pj = sx[v];
                    const float4 vj = sv[v];
                    const float pjv[3] = {pj.x, pj.y, pj.z};
                    const float vjv[3] = {vj.x, vj.y, vj.z};
                    float dr[D], dv[D];
                    float r2 = 0.f, drdv = 0.f;
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        dr[d] = xi[d] - pjv[d];
                        r2 += dr[d] * dr[d];
                    }
                    if (r2 <= 0.f) continue;  // src/force.h:42
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        dv[d] = vi[d] - vjv[d];
                        drdv += dr[d] * dv[d];
                    }
                    float mu;
                    const float den = pair_mu(r2, drdv, I, pj.w, a.k, mu, vsig);  // max: a repeat is harmless
                    const float4 qj = sq[v];
                    JState<float> J{pj.w, qj.x, qj.y, qj.z, qj.w, vj.w};
                    pair_force<D>(dr, dv, r2, drdv, den, mu, I, J, a.k, acc, divv);
                }
            }
        }
        // combine the parts of each i in part order
        __syncthreads();  // the last tile's sv and sq are read no more
#pragma unroll
        for (int d = 0; d < D; ++d) red[d][t] = acc[d];
        red[D][t] = divv;
        red[D + 1][t] = vsig;
        __syncthreads();
        if (t < nr) {
            for (int p = 1; p < S; ++p) {
                const int o = t + p * nr;
#pragma unroll
                for (int d = 0; d < D; ++d) acc[d] += red[d][o];
                divv += red[D][o];
                vsig = fmaxf(vsig, red[D + 1][o]);
            }
            if (out < 0) {
                float* r = a.res + (a.i_base + i) * (D + 2);
#pragma unroll
                for (int d = 0; d < D; ++d) r[d] = acc[d];
                r[D] = a.c_cfl * I.h / (I.c + vsig);  // src/force.h:56
                r[D + 1] = divv / I.rho;              // src/force.h:57
            } else {
                float* r = a.part + out + (int64_t)(i - i0) * (D + 2);
#pragma unroll
                for (int d = 0; d < D; ++d) r[d] = acc[d];
                r[D] = divv;
                r[D + 1] = vsig;
            }
        }
    }
}

// Results of the walks cut into pieces: the pieces' partial sums of each i
// added in piece order (one workgroup per such walk).
template <int D>
__global__ __launch_bounds__(kWalkThreads) void k_walk_combine(WalkArgs<D> a) {
    const int s = blockIdx.x;
    const int w = a.sp_walk[s];
    const int np = a.sp_np[s];
    const int i0 = a.ioff[w], n_i = a.ioff[w + 1] - i0;
    const float* base = a.part + a.sp_base[s];
    for (int li = (int)threadIdx.x; li < n_i; li += kWalkThreads) {
        float acc[D];
#pragma unroll
        for (int d = 0; d < D; ++d) acc[d] = 0.f;
        float divv = 0.f, vsig = 0.f;
        for (int p = 0; p < np; ++p) {
            const float* q = base + ((int64_t)p * n_i + li) * (D + 2);
#pragma unroll
            for (int d = 0; d < D; ++d) acc[d] += q[d];
            divv += q[D];
            vsig = fmaxf(vsig, q[D + 1]);
        }
        const int i = i0 + li;
        const float h = a.iq[0 * a.ni + i], rho = a.iq[1 * a.ni + i], c = a.iq[3 * a.ni + i];
        float* r = a.res + (a.i_base + i) * (D + 2);
#pragma unroll
        for (int d = 0; d < D; ++d) r[d] = acc[d];
        r[D] = a.c_cfl * h / (c + vsig);  // src/force.h:56
        r[D + 1] = divv / rho;            // src/force.h:57
    }
}

constexpr int kMaxChunks = 8;

struct TagState {
    bool ready = false;
    bool outstanding = false;
    int dim = 0;
    hipStream_t stream = nullptr;       // kernels and the result copy
    hipStream_t copy_stream = nullptr;  // uploads, so chunk c + 1 moves while chunk c computes
    hipStream_t stream2 = nullptr;      // odd chunks' kernels, so a chunk's tail overlaps the next chunk
    hipEvent_t ev_k2 = nullptr;         // stream2's last kernel done
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
    float* d_part = nullptr;  // partial sums of the pieces of long walks
    size_t part_cap = 0;
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
    if (t.d_part) (void)hipFree(t.d_part);
    (void)hipStreamSynchronize(t.copy_stream);
    (void)hipEventDestroy(t.ev_begin);
    (void)hipEventDestroy(t.ev_end);
    for (auto& e : t.ev_k) (void)hipEventDestroy(e);
    for (auto& e : t.ev_up) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(t.stream);
    (void)hipStreamDestroy(t.copy_stream);
    (void)hipStreamSynchronize(t.stream2);
    (void)hipStreamDestroy(t.stream2);
    (void)hipEventDestroy(t.ev_k2);
    t = TagState{};
}

size_t align_up(size_t v, size_t a = 256) { return (v + a - 1) / a * a; }

// Byte layout of one staged chunk (identical on host and device).
struct Layout {
    size_t ioff, joff, it_walk, it_jb, it_je, it_out, sp_walk, sp_base, sp_np, ix, iv, iq, jx, jv, jq, total;
};

// The work items of one chunk: every walk, the long ones cut along their j
// list into pieces of about kWalkPiece serial candidates per lane.
struct ChunkPlan {
    std::vector<int32_t> it_walk, it_jb, it_je, it_out;
    std::vector<int32_t> sp_walk, sp_base, sp_np;
    int64_t part_floats = 0;
};

Layout layout(int D, int n_walk, int64_t ni, int64_t nj, const ChunkPlan& P) {
    Layout L{};
    size_t o = 0;
    const size_t nit = std::max<size_t>(P.it_walk.size(), 1), nsp = std::max<size_t>(P.sp_walk.size(), 1);
    L.ioff = o; o = align_up(o + sizeof(int32_t) * (n_walk + 1));
    L.joff = o; o = align_up(o + sizeof(int32_t) * (n_walk + 1));
    L.it_walk = o; o = align_up(o + sizeof(int32_t) * nit);
    L.it_jb = o; o = align_up(o + sizeof(int32_t) * nit);
    L.it_je = o; o = align_up(o + sizeof(int32_t) * nit);
    L.it_out = o; o = align_up(o + sizeof(int32_t) * nit);
    L.sp_walk = o; o = align_up(o + sizeof(int32_t) * nsp);
    L.sp_base = o; o = align_up(o + sizeof(int32_t) * nsp);
    L.sp_np = o; o = align_up(o + sizeof(int32_t) * nsp);
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

// Work items of walks [w0, w1). A wave's time is its lanes' serial
// candidate count, the sum over i rounds of ceil(n_j / parts); a walk whose
// count exceeds kWalkPiece is cut into that many equal j ranges, each its
// own workgroup writing partial sums, so the longest walks no longer set
// the chunk's time. Items are launched longest first: a workgroup
// dispatched early is older, and the oldest waves of a SIMD issue first.
ChunkPlan plan_chunk(int D, int w0, int w1, const std::vector<int32_t>& ioff, const std::vector<int32_t>& joff) {
    ChunkPlan P;
    std::vector<std::pair<int64_t, int32_t>> cost;  // (-cost, item)
    for (int w = w0; w < w1; ++w) {
        const int lw = w - w0;
        const int n_i = ioff[w + 1] - ioff[w], n_j = joff[w + 1] - joff[w];
        const int jw = joff[w] - joff[w0];
        if (n_i == 0) continue;
        int64_t c = 0;
        for (int r = 0; r < n_i; r += kWalkThreads) {
            const int parts = kWalkThreads / std::min(kWalkThreads, n_i - r);
            c += (n_j + parts - 1) / parts;
        }
        int np = (int)std::min<int64_t>((c + kWalkPiece - 1) / kWalkPiece, std::max(1, n_j / kWalkMinPieceJ));
        np = std::max(np, 1);
        int32_t out = -1;
        if (np > 1) {
            WCSPH_REQUIRE(P.part_floats + (int64_t)np * n_i * (D + 2) < INT32_MAX, WCSPH_EOVERFLOW,
                          "partial sums exceed 2^31 floats");
            out = (int32_t)P.part_floats;
            P.sp_walk.push_back(lw);
            P.sp_base.push_back(out);
            P.sp_np.push_back(np);
            P.part_floats += (int64_t)np * n_i * (D + 2);
        }
        for (int p = 0; p < np; ++p) {
            cost.push_back({-(c / np), (int32_t)P.it_walk.size()});
            P.it_walk.push_back(lw);
            P.it_jb.push_back(jw + (int32_t)((int64_t)n_j * p / np));
            P.it_je.push_back(jw + (int32_t)((int64_t)n_j * (p + 1) / np));
            P.it_out.push_back(np > 1 ? out + p * n_i * (D + 2) : -1);
        }
    }
    std::sort(cost.begin(), cost.end());
    ChunkPlan S = P;
    for (size_t k = 0; k < cost.size(); ++k) {
        const int32_t src = cost[k].second;
        S.it_walk[k] = P.it_walk[src];
        S.it_jb[k] = P.it_jb[src];
        S.it_je[k] = P.it_je[src];
        S.it_out[k] = P.it_out[src];
    }
    return S;
}

// Pack walks [w0, w1) of the batch into one chunk block at h (SoA, FP32,
// positions and velocities relative to each walk's first i particle).
template <int D, class EPI, class EPJ>
void pack_chunk(char* h, const Layout& L, const ChunkPlan& P, int w0, int w1, const std::vector<int32_t>& ioff,
                const std::vector<int32_t>& joff, const EPI* const* epi, const EPJ* const* epj) {
    const int nw = w1 - w0;
    const int64_t ni = ioff[w1] - ioff[w0], nj = joff[w1] - joff[w0];
    int32_t* lio = (int32_t*)(h + L.ioff);
    int32_t* ljo = (int32_t*)(h + L.joff);
    for (int w = 0; w <= nw; ++w) {
        lio[w] = ioff[w0 + w] - ioff[w0];
        ljo[w] = joff[w0 + w] - joff[w0];
    }
    auto put = [h](size_t at, const std::vector<int32_t>& v) {
        if (!v.empty()) std::memcpy(h + at, v.data(), sizeof(int32_t) * v.size());
    };
    put(L.it_walk, P.it_walk);
    put(L.it_jb, P.it_jb);
    put(L.it_je, P.it_je);
    put(L.it_out, P.it_out);
    put(L.sp_walk, P.sp_walk);
    put(L.sp_base, P.sp_base);
    put(L.sp_np, P.sp_np);
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
        WCSPH_HIP(hipStreamCreateWithFlags(&T.stream2, hipStreamNonBlocking));
        WCSPH_HIP(hipEventCreateWithFlags(&T.ev_k2, hipEventDisableTiming));
        WCSPH_HIP(hipEventCreate(&T.ev_begin));
        WCSPH_HIP(hipEventCreate(&T.ev_end));
        for (auto& e : T.ev_k) WCSPH_HIP(hipEventCreate(&e));
        for (auto& e : T.ev_up) WCSPH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        T.ready = true;
    }
    WCSPH_HIP(hipSetDevice(B.p.device));
    // uploads and kernels of a dispatch that failed part way may still read
    // the staging buffers: let them finish before those are packed again
    WCSPH_HIP(hipStreamSynchronize(T.copy_stream));
    WCSPH_HIP(hipStreamSynchronize(T.stream));
    WCSPH_HIP(hipStreamSynchronize(T.stream2));
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
    std::vector<ChunkPlan> plans(nch);
    std::vector<Layout> Ls(nch);
    std::vector<size_t> base(nch + 1, 0);
    std::vector<int64_t> pbase(nch + 1, 0);
    for (int c = 0; c < nch; ++c) {
        plans[c] = plan_chunk(D, wb[c], wb[c + 1], ioff, joff);
        Ls[c] = layout(D, wb[c + 1] - wb[c], ioff[wb[c + 1]] - ioff[wb[c]], joff[wb[c + 1]] - joff[wb[c]],
                       plans[c]);
        base[c + 1] = base[c] + Ls[c].total;
        pbase[c + 1] = pbase[c] + align_up(plans[c].part_floats * sizeof(float)) / sizeof(float);
    }
    const size_t part_bytes = std::max<size_t>(sizeof(float) * pbase[nch], 256);
    if (part_bytes > T.part_cap) {
        if (T.d_part) WCSPH_HIP(hipFree(T.d_part));
        T.d_part = nullptr;
        const size_t cap = part_bytes + part_bytes / 4;
        WCSPH_HIP(hipMalloc((void**)&T.d_part, cap));
        T.part_cap = cap;
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
    a.k.grad_mode = B.p.grad_mode;
    a.c_cfl = (float)B.p.c_cfl;
    WCSPH_HIP(hipEventRecord(T.ev_begin, T.copy_stream));
    double pack_ms = 0.0;
    for (int c = 0; c < nch; ++c) {
        const int w0 = wb[c], w1 = wb[c + 1];
        const Layout& L = Ls[c];
        char* h = T.h_in + base[c];
        const auto t0 = std::chrono::steady_clock::now();
        pack_chunk<D>(h, L, plans[c], w0, w1, ioff, joff, epi, epj);
        pack_ms += ms_since(t0);
        // the chunk goes to the device while the next one is packed, and its
        // kernel waits only for its own upload
        WCSPH_HIP(hipMemcpyAsync(T.d_in + base[c], h, L.total, hipMemcpyHostToDevice, T.copy_stream));
        WCSPH_HIP(hipEventRecord(T.ev_up[c], T.copy_stream));
        // even chunks' kernels on the tag's stream, odd ones on stream2
        hipStream_t ks = (c & 1) ? T.stream2 : T.stream;
        WCSPH_HIP(hipStreamWaitEvent(ks, T.ev_up[c], 0));
        char* d = T.d_in + base[c];
        a.ioff = (const int32_t*)(d + L.ioff);
        a.joff = (const int32_t*)(d + L.joff);
        a.it_walk = (const int32_t*)(d + L.it_walk);
        a.it_jb = (const int32_t*)(d + L.it_jb);
        a.it_je = (const int32_t*)(d + L.it_je);
        a.it_out = (const int32_t*)(d + L.it_out);
        a.sp_walk = (const int32_t*)(d + L.sp_walk);
        a.sp_base = (const int32_t*)(d + L.sp_base);
        a.sp_np = (const int32_t*)(d + L.sp_np);
        a.part = T.d_part + pbase[c];
        a.ix = (const float*)(d + L.ix);
        a.iv = (const float*)(d + L.iv);
        a.iq = (const float*)(d + L.iq);
        a.jx = (const float*)(d + L.jx);
        a.jv = (const float*)(d + L.jv);
        a.jq = (const float*)(d + L.jq);
        a.ni = ioff[w1] - ioff[w0];
        a.nj = joff[w1] - joff[w0];
        a.i_base = ioff[w0];
        WCSPH_HIP(hipEventRecord(T.ev_k[2 * c], ks));
        const int nit = (int)plans[c].it_walk.size(), nsp = (int)plans[c].sp_walk.size();
        if (nit > 0) {
            if (B.p.vsig_mode)
                hipLaunchKernelGGL((k_walk_force<D, 1>), dim3(nit), dim3(kWalkThreads), 0, ks, a);
            else
                hipLaunchKernelGGL((k_walk_force<D, 0>), dim3(nit), dim3(kWalkThreads), 0, ks, a);
        }
        if (nsp > 0) hipLaunchKernelGGL((k_walk_combine<D>), dim3(nsp), dim3(kWalkThreads), 0, ks, a);
        WCSPH_HIP(hipGetLastError());
        WCSPH_HIP(hipEventRecord(T.ev_k[2 * c + 1], ks));
    }
    // the result copy follows the kernels of both streams
    WCSPH_HIP(hipEventRecord(T.ev_k2, T.stream2));
    WCSPH_HIP(hipStreamWaitEvent(T.stream, T.ev_k2, 0));
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
        // kernel time: the union of the chunk kernels' intervals (kernels of
        // consecutive chunks may overlap on the two streams)
        std::vector<std::pair<float, float>> iv((size_t)T.nchunks);
        for (int c = 0; c < T.nchunks; ++c) {
            WCSPH_HIP(hipEventElapsedTime(&iv[c].first, T.ev_begin, T.ev_k[2 * c]));
            WCSPH_HIP(hipEventElapsedTime(&iv[c].second, T.ev_begin, T.ev_k[2 * c + 1]));
        }
        std::sort(iv.begin(), iv.end());
        float k = 0.f, cur_a = 0.f, cur_b = -1.f;
        for (const auto& x : iv) {
            if (x.first > cur_b) {
                if (cur_b > cur_a) k += cur_b - cur_a;
                cur_a = x.first;
                cur_b = x.second;
            } else {
                cur_b = std::max(cur_b, x.second);
            }
        }
        if (cur_b > cur_a) k += cur_b - cur_a;
        float ms = 0.f;
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
        WCSPH_REQUIRE(p->grad_mode == 0 || p->grad_mode == 1, WCSPH_EINVAL, "grad_mode must be 0 or 1");
        WCSPH_REQUIRE(p->wall_mode == 0, WCSPH_EINVAL,
                      "the FDPS drop-in evaluates every pair (wall_mode 0); wall_mode 1 is for the device-resident step");
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
