// shim_demo.cpp -- a reference-style caller compiled against the drop-in
// headers (include/gsap/*.h): the DemoParticles::step order with Sap,
// ParticleRigidBody and RadixSort used by their reference names.
// Usage: shim_demo <in.bin> <out.bin>
//   in.bin : int32 n, then n float4 spheres (x y z r)
//   out.bin: int32 NPairs, int32 NPEntries, NPairs int2, NPEntries int,
//            then n uint32 keys sorted by RadixSort and their values
#include <cmath>
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

    // The same collision inside a caller-provided workspace: identical output.
    {
        Sap ws_sap(n);
        ws_sap.setPairCapacity(maxPairs);
        const unsigned long long wsb = ws_sap.workspaceSize();
        void* ws = nullptr;
        HC(hipMalloc(&ws, wsb));
        ws_sap.setWorkspace(ws, wsb);
        if (ws_sap.lastStatus() != GSAP_OK) return 6;
        int2* p2 = nullptr;
        int* e2 = nullptr;
        HC(hipMalloc(&p2, 8ull * maxPairs));
        HC(hipMalloc(&e2, 4ull * n));
        int np2 = 0, ne2 = 0;
        ws_sap.collision(1.0f, d, p2, e2, &np2, &ne2);
        if (ws_sap.lastStatus() != GSAP_OK || np2 != np || ne2 != ne) return 7;
        std::vector<int2> hp2(np2);
        std::vector<int> he2(ne2);
        HC(hipMemcpy(hp2.data(), p2, 8ull * np2, hipMemcpyDeviceToHost));
        HC(hipMemcpy(he2.data(), e2, 4ull * ne2, hipMemcpyDeviceToHost));
        for (int k = 0; k < np; ++k)
            if (hp2[k].x != hp[k].x || hp2[k].y != hp[k].y) return 8;
        for (int k = 0; k < ne; ++k)
            if (he2[k] != he[k]) return 8;
        (void)hipFree(p2); (void)hipFree(e2); (void)hipFree(ws);
    }

    // ParticleRigidBody set up the way DemoParticles::init does
    // (demoParticles.h:230-414): host RigidBodyData with Quaternion and
    // Matrix3x3 arrays, templates in MAX_NUM_PARTICLES slots per shape, shape
    // inverse inertias from Matrix3x3::invert, setParams / setMaxVelocity.
    {
        const int nb = 2;
        ParticleRigidBody prb(nb, 2 * 3, 64);
        float4 pos[nb] = {{0, 5, 0, 0}, {1.9f, 5, 0, 0}};
        Quaternion q[nb] = {Quaternion(0, 0, 0, 1), Quaternion(0, 0, 0.38268343f, 0.92387953f)};
        float4 lv[nb] = {{1, 0, 0, 0}, {-1, 0, 0, 0}}, av[nb] = {{0, 0, 0, 0}, {0, 0.5f, 0, 0}};
        int shape[nb] = {0, 1}, start[nb] = {0, 3}, cnt[nb] = {3, 3};
        float im[nb] = {1.0f / 3, 1.0f / 6};
        Matrix3x3 ii[nb];
        for (int b = 0; b < nb; ++b) ii[b].setIdentity();
        ParticleRigidBody::RigidBodyData rb = {nb, pos, q, lv, av, shape, start, cnt, im, ii, 6};
        prb.setRbData(rb);
        float4 ref[2 * MAX_NUM_PARTICLES] = {};
        float mass[2 * MAX_NUM_PARTICLES] = {};
        for (int s = 0; s < 2; ++s)
            for (int k = 0; k < 3; ++k) {
                ref[ParticleRigidBody::getParticleRefIdx(s, k)] = make_float4(0.5f * (k - 1), 0.1f * s, 0, 0.25f);
                mass[ParticleRigidBody::getParticleRefIdx(s, k)] = 1.0f + s;
            }
        prb.setParticleData(ref, mass, 2 * MAX_NUM_PARTICLES);
        Matrix3x3 shapeInv[2] = {Matrix3x3(make_float4(2, 0, 0, 0), make_float4(0, 3, 0, 0), make_float4(0, 0, 4, 0)),
                                 Matrix3x3(make_float4(4, 1, 0, 0), make_float4(0, 2, 0, 0), make_float4(0, 0, 1, 0))};
        for (int s = 0; s < 2; ++s) shapeInv[s].invert();
        prb.setParticleShapeInvInertia(shapeInv, 2);
        prb.setParams(2.1f, 0.8f);
        prb.setMaxVelocity(0.25f, 0.01f);

        // getRbShapeIdx / getRbNumParticles as demoParticles.h:1398-1401 reads them
        int hs[nb], hn[nb];
        HC(hipMemcpy(hs, prb.getRbShapeIdx(), sizeof hs, hipMemcpyDeviceToHost));
        HC(hipMemcpy(hn, prb.getRbNumParticles(), sizeof hn, hipMemcpyDeviceToHost));
        for (int b = 0; b < nb; ++b)
            if (hs[b] != shape[b] || hn[b] != cnt[b]) return 9;

        // particle positions = x + q r q^-1 (the header's Quaternion on the host)
        prb.updateParticleData();
        std::vector<float4> pp(6);
        std::vector<Quaternion> gq(nb);
        HC(hipMemcpy(pp.data(), prb.getParticlePos(), 6 * 16, hipMemcpyDeviceToHost));
        HC(hipMemcpy(gq.data(), prb.getRbQuat(), nb * 16, hipMemcpyDeviceToHost));
        for (int b = 0; b < nb; ++b)
            for (int k = 0; k < 3; ++k) {
                float4 r;
                gq[b].rotateVector(ref[ParticleRigidBody::getParticleRefIdx(shape[b], k)], r);
                const float4 g = pp[start[b] + k];
                if (std::fabs(g.x - (r.x + pos[b].x)) > 1e-5f || std::fabs(g.y - (r.y + pos[b].y)) > 1e-5f ||
                    std::fabs(g.z - (r.z + pos[b].z)) > 1e-5f || g.w != 0.25f)
                    return 10;
            }
        // prepare: I_w^-1 = R I0^-1 R^T with R from the header's quaternion
        prb.prepare();
        gsap_prb_ptrs ptrs;
        if (gsap_prb_get_ptrs(prb.handle(), &ptrs) != GSAP_OK) return 11;
        Matrix3x3 iw[nb];
        HC(hipMemcpy(iw, ptrs.rb_inv_inertia, sizeof iw, hipMemcpyDeviceToHost));
        for (int b = 0; b < nb; ++b) {
            Matrix3x3 R, RT;
            gq[b].getRotationMatrix(R);
            RT.setTranspose(R);
            const Matrix3x3 want = (R * shapeInv[shape[b]]) * RT;
            for (int i = 0; i < 3; ++i)
                if (std::fabs(iw[b][i].x - want[i].x) > 1e-5f || std::fabs(iw[b][i].y - want[i].y) > 1e-5f ||
                    std::fabs(iw[b][i].z - want[i].z) > 1e-5f)
                    return 11;
        }

        Sap s2(6);
        for (int k = 0; k < 3; ++k) prb.step(s2, 0.001f);
        std::vector<float4> v(nb);
        HC(hipMemcpy(v.data(), prb.getRbPos(), nb * 16, hipMemcpyDeviceToHost));
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
