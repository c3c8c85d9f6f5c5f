// Drives GPU_WC::HydrDispatchKernel / HydrRetrieveKernel the way FDPS's
// calcForceAllAndWriteBackMultiWalk does: for each batch of at most
// n_walk_limit walks, retrieve(previous batch) then dispatch(this batch),
// and a final retrieve. Input and output are raw record files written and
// read by tests/test_gpu_fdps.py.
//
// usage: bridge_driver in.bin out.bin n_walk_limit
// in.bin : int32 n_walk, int32 n_epi[n_walk], int32 n_epj[n_walk],
//          EPI records (all walks), EPJ records (all walks),
//          then the wcsph_params struct
// out.bin: RESULT records (all walks)
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mock_fdps_types.hpp"
#include "wcsph_fdps.hpp"

template <class T>
static bool rd(FILE* f, T* p, size_t n) { return fread(p, sizeof(T), n, f) == n; }

int main(int argc, char** argv) {
    if (argc != 4) return 2;
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 2;
    int32_t nw = 0;
    if (!rd(f, &nw, 1)) return 2;
    std::vector<int32_t> ni(nw), nj(nw);
    if (!rd(f, ni.data(), nw) || !rd(f, nj.data(), nw)) return 2;
    size_t ti = 0, tj = 0;
    for (int w = 0; w < nw; ++w) { ti += ni[w]; tj += nj[w]; }
    std::vector<STD::EPI::Hydro> epi(ti);
    std::vector<STD::EPJ::Hydro> epj(tj);
    wcsph_params p;
    if (!rd(f, epi.data(), ti) || !rd(f, epj.data(), tj) || !rd(f, &p, 1)) return 2;
    fclose(f);
    if (GPU_WC::Init(p) != WCSPH_OK) { fprintf(stderr, "%s\n", wcsph_last_error()); return 3; }
    std::vector<STD::RESULT::Hydro> res(ti);
    std::vector<const STD::EPI::Hydro*> pi(nw);
    std::vector<const STD::EPJ::Hydro*> pj(nw);
    std::vector<STD::RESULT::Hydro*> pr(nw);
    size_t oi = 0, oj = 0;
    for (int w = 0; w < nw; ++w) {
        pi[w] = epi.data() + oi; pr[w] = res.data() + oi; pj[w] = epj.data() + oj;
        oi += ni[w]; oj += nj[w];
    }
    const int limit = atoi(argv[3]);
    int prev_b = -1, prev_n = 0;
    for (int b = 0; b < nw; b += limit) {
        const int n = nw - b < limit ? nw - b : limit;
        if (prev_b >= 0 && GPU_WC::HydrRetrieveKernel(0, prev_n, &ni[prev_b], &pr[prev_b]) != 0) {
            fprintf(stderr, "%s\n", wcsph_last_error()); return 4;
        }
        if (GPU_WC::HydrDispatchKernel(0, n, &pi[b], &ni[b], &pj[b], &nj[b]) != 0) {
            fprintf(stderr, "%s\n", wcsph_last_error()); return 4;
        }
        prev_b = b; prev_n = n;
    }
    if (prev_b >= 0 && GPU_WC::HydrRetrieveKernel(0, prev_n, &ni[prev_b], &pr[prev_b]) != 0) {
        fprintf(stderr, "%s\n", wcsph_last_error()); return 4;
    }
    FILE* g = fopen(argv[2], "wb");
    if (!g) return 2;
    fwrite(res.data(), sizeof(STD::RESULT::Hydro), ti, g);
    fclose(g);
    wcsph_finalize();
    return 0;
}
