// shim_demo.cpp -- a reference-style caller compiled against the drop-in
// headers (include/gsap/*.h): the DemoParticles::step order with Sap,
// ParticleRigidBody and RadixSort used by their reference names.
// Usage: shim_demo <in.bin> <out.bin>
//   in.bin : int32 n, then n float4 spheres (x y z r)
//   out.bin: int32 NPairs, int32 NPEntries, NPairs int2, NPEntries int,
//            then n uint32 keys sorted by RadixSort and their values
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "gsap/ParticleRigidBody.h"
#include "gsap/Sap.h"
#include "gsap/radixsort.h"

#define HC(x)                                                                  \
    do {                                                                       \
        if ((x) != hipSuccess) {                                               \
            std::fprintf(stderr, "%s failed\n", #x);                           \
            return 3;                                                          \
        }                                                                      \
    } while (0)

int main(int argc, char** argv) {
    if (argc != 3) return 2;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    int n = 0;
    if (std::fread(&n, 4, 1, f) != 1 || n <= 0) return 2;
    std::vector<float4> h(n);
    if (std::fread(h.data(), 16, n, f) != (size_t)n) return 2;
    std::fclose(f);

    float4* d = nullptr;
    int2* pairs = nullptr;
    int* entries = nullptr;
    const int maxPairs = 64 * n;
    HC(hipMalloc(&d, 16ull * n));
    HC(hipMalloc(&pairs, 8ull * maxPairs));
    HC(hipMalloc(&entries, 4ull * n));
    HC(hipMemcpy(d, h.data(), 16ull * n, hipMemcpyHostToDevice));

    Sap sap(n);
    sap.setPairCapacity(maxPairs);
    int np = 0, ne = 0;
    sap.collision(1.0f, d, pairs, entries, &np, &ne);
    if (sap.lastStatus() != GSAP_OK) return 4;
    std::vector<int2> hp(np);
    std::vector<int> he(ne);
    HC(hipMemcpy(hp.data(), pairs, 8ull * np, hipMemcpyDeviceToHost));
    HC(hipMemcpy(he.data(), entries, 4ull * ne, hipMemcpyDeviceToHost));

    // RadixSort on the flipped-free float x coordinates with index values
    std::vector<float> keys(n);
    std::vector<unsigned> vals(n);
    for (int i = 0; i < n; ++i) { keys[i] = h[i].x; vals[i] = (unsigned)i; }
    float* dk = nullptr;
    unsigned* dv = nullptr;
    HC(hipMalloc(&dk, 4ull * n));
    HC(hipMalloc(&dv, 4ull * n));
    HC(hipMemcpy(dk, keys.data(), 4ull * n, hipMemcpyHostToDevice));
    HC(hipMemcpy(dv, vals.data(), 4ull * n, hipMemcpyHostToDevice));
    {
        nvRadixSort::RadixSort rs(n);
        rs.sort(dk, dv, n, 32, true);
    }
    HC(hipDeviceSynchronize());
    HC(hipMemcpy(keys.data(), dk, 4ull * n, hipMemcpyDeviceToHost));
    HC(hipMemcpy(vals.data(), dv, 4ull * n, hipMemcpyDeviceToHost));

    // ParticleRigidBody: two single-particle bodies, reference defaults
    {
        ParticleRigidBody prb(2, 2, 16);
        float4 pos[2] = {{0, 5, 0, 0}, {0.9f, 5, 0, 0}}, q[2] = {{0, 0, 0, 1}, {0, 0, 0, 1}};
        float4 lv[2] = {{1, 0, 0, 0}, {-1, 0, 0, 0}}, av[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
        int shape[2] = {0, 0}, start[2] = {0, 1}, cnt[2] = {1, 1};
        float im[2] = {1, 1};
        ParticleRigidBody::RigidBodyData rb = {2, pos, q, lv, av, shape, start, cnt, im, nullptr, 2};
        prb.setRbData(rb);
        float4 ref = {0, 0, 0, 0.5f};
        float mass = 1.0f;
        prb.setParticleData(&ref, &mass, 1, 1);
        Sap s2(2);
        for (int k = 0; k < 3; ++k) prb.step(s2, 0.001f);
        std::vector<float4> v(2);
        HC(hipMemcpy(v.data(), prb.getRbPos(), 32, hipMemcpyDeviceToHost));
        std::fprintf(stderr, "prb pos %f %f\n", v[0].x, v[1].x);
        if (!(v[0].x < v[1].x)) return 5;
    }

    FILE* o = std::fopen(argv[2], "wb");
    if (!o) return 2;
    std::fwrite(&np, 4, 1, o);
    std::fwrite(&ne, 4, 1, o);
    std::fwrite(hp.data(), 8, np, o);
    std::fwrite(he.data(), 4, ne, o);
    std::fwrite(keys.data(), 4, n, o);
    std::fwrite(vals.data(), 4, n, o);
    std::fclose(o);
    (void)hipFree(d); (void)hipFree(pairs); (void)hipFree(entries); (void)hipFree(dk); (void)hipFree(dv);
    return 0;
}
