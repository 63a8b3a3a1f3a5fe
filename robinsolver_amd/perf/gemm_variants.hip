// gemm_variants.hip -- times K-loop variants of the batched complex GEMM
// (robinsolver_amd/csrc/k_gemm.hip) on solve-pass shapes and checks that every
// variant gives the bits of the first.  Not part of the library.
//   (no argument)  the 2D 1024^2 tree's shapes, N = 256 right-hand sides
//   3d             the 3D 64^3 tree's top-level shapes, N = 16
//
//   hipcc -x hip --offload-arch=gfx950 -Xclang -target-feature -Xclang -packed-fp32-ops \
//         -mllvm -amdgpu-mfma-vgpr-form -O3 -std=c++17 \
//         robinsolver_amd/perf/gemm_variants.hip -o gemm_variants && ./gemm_variants [3d]
// (the library's own flags, robinsolver_amd/build.py)
#include "../csrc/k_gemm.hip"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

using namespace rs;

#define CHK(x)                                                                      \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                     \
            std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));     \
            std::exit(1);                                                           \
        }                                                                           \
    } while (0)

template <int BM, int BN, int BK, bool DB>
void launch_var(const GemmDesc* d, int np, int M, int N, int K, void* ws) {
    int ks = gemm_kslices(np, M, K, solve_ks_target(np, M, K));
    const int kchunk = ((K + ks - 1) / ks + 31) / 32 * 32;
    ks = (K + kchunk - 1) / kchunk;
    const int tilesM = (M + BM - 1) / BM, tilesN = (N + BN - 1) / BN;
    hipLaunchKernelGGL((gemm_kernel<float, float, BM, BN, BK, DB>), dim3(np * tilesM * tilesN * ks), dim3(256), 0,
                       nullptr, d, M, N, K, tilesM, tilesN, ks, kchunk, (float2*)ws);
    if (ks > 1) {
        const size_t mn = (size_t)M * N;
        hipLaunchKernelGGL((gemm_reduce_kernel<float, float>), dim3((unsigned)((mn + 255) / 256), np), dim3(256), 0,
                           nullptr, d, M, N, ks, (const float2*)ws);
    }
}

struct Var {
    const char* name;
    void (*f)(const GemmDesc*, int, int, int, int, void*);
};

struct Shape {
    const char* name;
    int M, K, np;
    bool din;
};

int main(int argc, char** argv) {
    const bool d3 = argc > 1 && std::strcmp(argv[1], "3d") == 0;
    // 3D 64^3, 8^3 leaves: level 0 splits x (i = 4096), level 1 y (2048), level 2 z (1024)
    const Shape shapes3[] = {{"upg_l0", 8192, 8192, 1, false}, {"dnt_l0", 4096, 12288, 2, false},
                             {"uph_l0", 12288, 4096, 2, true}, {"upg_l1", 4096, 4096, 2, false},
                             {"dnt_l1", 2048, 8192, 4, false}, {"upg_l2", 2048, 2048, 4, false},
                             {"dnt_l2", 1024, 6144, 8, false}, {"upg_l3", 2048, 2048, 8, false}};
    const Var vars3[] = {{"64x16 bk16 2buf", launch_var<64, 16, 16, true>},
                         {"64x16 bk32 2buf", launch_var<64, 16, 32, true>},
                         {"64x16 bk64 2buf", launch_var<64, 16, 64, true>},
                         {"64x16 bk32 1buf", launch_var<64, 16, 32, false>},
                         {"64x16 bk64 1buf", launch_var<64, 16, 64, false>},
                         {"128x16 bk16 2buf", launch_var<128, 16, 16, true>},
                         {"128x16 bk32 2buf", launch_var<128, 16, 32, true>},
                         {"128x16 bk64 1buf", launch_var<128, 16, 64, false>},
                         {"32x32 bk16 2buf", launch_var<32, 32, 16, true>},
                         {"64x64 bk16 2buf", launch_var<64, 64, 16, true>}};
    const Shape shapes2[] = {{"upg_l0", 2048, 2048, 1, false}, {"upg_l1", 1024, 1024, 2, false},
                  {"upg_l2", 1024, 1024, 4, false}, {"upg_l4", 512, 512, 16, false},
                  {"upg_l6", 256, 256, 64, false},  {"upg_l8", 128, 128, 256, false},
                  {"upg_l10", 64, 64, 1024, false}, {"uph_l2", 1024, 512, 8, true},
                  {"dnt_l2", 512, 1024, 8, false},  {"uph_l6", 256, 128, 128, true}};
    const Var vars2[] = {{"64x64 bk16 2buf", launch_var<64, 64, 16, true>},
                        {"64x64 bk32 1buf", launch_var<64, 64, 32, false>},
                        {"64x64 bk32 2buf", launch_var<64, 64, 32, true>},
                        {"64x64 bk64 1buf", launch_var<64, 64, 64, false>},
                        {"64x32 bk16 2buf", launch_var<64, 32, 16, true>},
                        {"64x32 bk32 1buf", launch_var<64, 32, 32, false>},
                        {"32x64 bk16 2buf", launch_var<32, 64, 16, true>},
                        {"32x32 bk16 2buf", launch_var<32, 32, 16, true>},
                        {"32x32 bk32 1buf", launch_var<32, 32, 32, false>},
                        {"64x16 bk16 2buf", launch_var<64, 16, 16, true>},
                        {"16x64 bk16 2buf", launch_var<16, 64, 16, true>},
                        {"128x16 bk16 2buf", launch_var<128, 16, 16, true>}};
    const Shape* shapes = d3 ? shapes3 : shapes2;
    const Var* vars = d3 ? vars3 : vars2;
    const int ns = d3 ? (int)(sizeof(shapes3) / sizeof(shapes3[0])) : (int)(sizeof(shapes2) / sizeof(shapes2[0]));
    const int nv = d3 ? (int)(sizeof(vars3) / sizeof(vars3[0])) : (int)(sizeof(vars2) / sizeof(vars2[0]));
    const int N = d3 ? 16 : 256, REPS = 20;
    hipEvent_t e0, e1;
    CHK(hipEventCreate(&e0));
    CHK(hipEventCreate(&e1));
    std::printf("| variant |");
    for (int si = 0; si < ns; ++si) std::printf(" %s |", shapes[si].name);
    std::printf(" sum |\n|---|");
    for (int i = 0; i <= ns; ++i) std::printf("---:|");
    std::printf("\n");
    std::vector<std::vector<double>> us(nv, std::vector<double>(ns));
    std::vector<std::vector<int>> same(nv, std::vector<int>(ns, 1));
    for (int si = 0; si < ns; ++si) {
        const Shape& sh = shapes[si];
        const int M = sh.M, K = sh.K, np = sh.np;
        std::vector<float2> hA((size_t)np * M * K), hB((size_t)np * K * N);
        for (size_t i = 0; i < hA.size(); ++i)
            hA[i] = make_float2((float)((i * 7919) % 1000) / 1000.f - 0.5f, (float)((i * 104729) % 997) / 997.f - 0.5f);
        for (size_t i = 0; i < hB.size(); ++i)
            hB[i] = make_float2((float)((i * 31) % 1013) / 1013.f - 0.5f, (float)((i * 17) % 1009) / 1009.f - 0.5f);
        float2 *A, *B, *C, *ws = nullptr;
        const size_t nc = (size_t)np * M * N;
        CHK(hipMalloc(&A, hA.size() * sizeof(float2)));
        CHK(hipMalloc(&B, hB.size() * sizeof(float2)));
        CHK(hipMalloc(&C, nc * sizeof(float2)));
        const size_t wse = gemm_ws_elems(np, M, N, K, solve_ks_target(np, M, K));
        if (wse) CHK(hipMalloc(&ws, wse * sizeof(float2)));
        CHK(hipMemcpy(A, hA.data(), hA.size() * sizeof(float2), hipMemcpyHostToDevice));
        CHK(hipMemcpy(B, hB.data(), hB.size() * sizeof(float2), hipMemcpyHostToDevice));
        std::vector<GemmDesc> hd(np);
        for (int p = 0; p < np; ++p) {
            GemmDesc& d = hd[p];
            std::memset(&d, 0, sizeof d);
            d.A = A + (size_t)p * M * K;
            d.lda = M;
            d.B = B + (size_t)p * K * N;
            d.ldb = K;
            d.ksplit = 1 << 30;
            d.Out = C + (size_t)p * M * N;
            d.ldo = M;
            if (sh.din) {   // beta * Out added in place (the timed launches accumulate)
                d.Din = d.Out;
                d.ldd = M;
            }
            d.ar = 1;
            d.br = 1;
        }
        GemmDesc* dd;
        CHK(hipMalloc(&dd, np * sizeof(GemmDesc)));
        CHK(hipMemcpy(dd, hd.data(), np * sizeof(GemmDesc), hipMemcpyHostToDevice));
        std::vector<float2> ref, out(nc);
        for (int vi = 0; vi < nv; ++vi) {
            const Var& v = vars[vi];
            for (int w = 0; w < 3; ++w) v.f(dd, np, M, N, K, ws);
            CHK(hipEventRecord(e0, nullptr));
            for (int r = 0; r < REPS; ++r) v.f(dd, np, M, N, K, ws);
            CHK(hipEventRecord(e1, nullptr));
            CHK(hipEventSynchronize(e1));
            float ms = 0.f;
            CHK(hipEventElapsedTime(&ms, e0, e1));
            us[vi][si] = 1e3 * ms / REPS;
            // one launch on a zeroed output: the bits of the product (plus zero)
            CHK(hipMemset(C, 0, nc * sizeof(float2)));
            v.f(dd, np, M, N, K, ws);
            CHK(hipMemcpy(out.data(), C, nc * sizeof(float2), hipMemcpyDeviceToHost));
            if (ref.empty()) ref = out;
            else same[vi][si] = std::memcmp(ref.data(), out.data(), nc * sizeof(float2)) == 0;
        }
        CHK(hipFree(A));
        CHK(hipFree(B));
        CHK(hipFree(C));
        CHK(hipFree(dd));
        if (ws) CHK(hipFree(ws));
        std::fprintf(stderr, "shape %s done\n", sh.name);
    }
    for (int vi = 0; vi < nv; ++vi) {
        double tot = 0;
        std::printf("| %s |", vars[vi].name);
        for (int si = 0; si < ns; ++si) {
            std::printf(" %.1f%s |", us[vi][si], same[vi][si] ? "" : " (bits differ)");
            tot += us[vi][si];
        }
        std::printf(" %.0f |\n", tot);
    }
    return 0;
}
