// shim_driver -- runs VertexSmoother / LocalOperations (the MI355X mirrors)
// on a mesh file and writes the result, so that a caller written against the
// reference's class surface can be exercised end to end.
//
//   shim_driver <mesh.bin> <out.bin> <mode>
//   mode: seq   smooth() in vertex order (the reference's Gauss-Seidel order)
//         batch smoothSingleBatched() + smoothSurfaceBatched()
//
// mesh.bin (little-endian): int64 nv, nt, nsf, nbf, is_mesh_closed;
// double eps, eps_2, sampling_dist; double posf[nv][3]; uint8 flags[nv]
// (1 surface, 2 bbox, 4 boundary, 8 locked); int32 tets[nt][4];
// int32 is_surface_fs[nt][4]; double tet_quality[nt];
// double sf_tri[nsf][9]; double b_tri[nbf][9].
// out.bin: double posf[nv][3]; double slim_energy[nt]; int64 counter,
// suc_counter; plus, for the first up-to-8 interior vertices, NewtonsUpdate
// (E, J[3], H[9]) and getNewEnergy at the final positions, and the static
// comformalAMIPSEnergy_new of tet 0.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "tetwild_shim.h"

using namespace tetwild_amd;

template <class T>
static bool rd(FILE* f, T* p, size_t n)
{
    return fread(p, sizeof(T), n, f) == n;
}

int main(int argc, char** argv)
{
    if (argc != 4) {
        fprintf(stderr, "usage: shim_driver mesh.bin out.bin seq|batch\n");
        return 2;
    }
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 2;
    int64_t hdr[5];
    double eps[3];
    if (!rd(f, hdr, 5) || !rd(f, eps, 3)) return 2;
    int64_t nv = hdr[0], nt = hdr[1], nsf = hdr[2], nbf = hdr[3];
    std::vector<TetVertex> tv(nv);
    std::vector<double> pos(3 * nv);
    std::vector<uint8_t> flags(nv);
    std::vector<std::array<int, 4>> tets(nt), fs(nt);
    std::vector<double> q(nt), sf(9 * nsf), bf(9 * nbf);
    if (!rd(f, pos.data(), 3 * nv) || !rd(f, flags.data(), nv) ||
        !rd(f, (int32_t*)tets.data(), 4 * nt) || !rd(f, (int32_t*)fs.data(), 4 * nt) ||
        !rd(f, q.data(), nt) || !rd(f, sf.data(), 9 * nsf) || !rd(f, bf.data(), 9 * nbf))
        return 2;
    fclose(f);
    for (int64_t v = 0; v < nv; ++v) {
        tv[v].posf = {{pos[3 * v], pos[3 * v + 1], pos[3 * v + 2]}};
        tv[v].is_on_surface = flags[v] & 1;
        tv[v].is_on_bbox = flags[v] & 2;
        tv[v].is_on_boundary = flags[v] & 4;
        tv[v].is_locked = flags[v] & 8;
    }
    for (int64_t t = 0; t < nt; ++t)
        for (int j = 0; j < 4; ++j) tv[tets[t][j]].conn_tets.insert((int)t);
    std::vector<TetQuality> tq(nt);
    for (int64_t t = 0; t < nt; ++t) tq[t].slim_energy = q[t];
    std::vector<bool> v_rm(nv, false), t_rm(nt, false);
    auto surface = [](const std::vector<double>& tri, std::vector<Point_3f>& V,
                      std::vector<std::array<int, 3>>& F) {
        size_t n = tri.size() / 9;
        for (size_t i = 0; i < n; ++i) {
            for (int k = 0; k < 3; ++k) V.push_back({{tri[9 * i + 3 * k], tri[9 * i + 3 * k + 1],
                                                      tri[9 * i + 3 * k + 2]}});
            F.push_back({{(int)(3 * i), (int)(3 * i + 1), (int)(3 * i + 2)}});
        }
    };
    std::vector<Point_3f> SV, BV;
    std::vector<std::array<int, 3>> SF, BF;
    surface(sf, SV, SF);
    surface(nbf ? bf : sf, BV, BF);
    try {
        Device dev(0);
        Envelope sf_tree(dev, SV, SF), b_tree(dev, BV, BF);
        State st;
        st.eps = eps[0];
        st.eps_2 = eps[1];
        st.sampling_dist = eps[2];
        st.is_mesh_closed = hdr[4] != 0;
        VertexSmoother vs(tv, tets, fs, v_rm, t_rm, tq, sf_tree, b_tree, st, dev);
        std::string mode = argv[3];
        int64_t counters[2];
        if (mode == "seq") {
            vs.smooth();
        } else {
            vs.smoothSingleBatched();
            if (st.eps >= 0) vs.smoothSurfaceBatched();
        }
        counters[0] = vs.counter;
        counters[1] = vs.suc_counter;
        FILE* o = fopen(argv[2], "wb");
        if (!o) return 2;
        for (int64_t v = 0; v < nv; ++v) fwrite(tv[v].posf.data(), 8, 3, o);
        for (int64_t t = 0; t < nt; ++t) fwrite(&tq[t].slim_energy, 8, 1, o);
        fwrite(counters, 8, 2, o);
        int done = 0;
        for (int64_t v = 0; v < nv && done < 8; ++v) {
            if (tv[v].is_on_surface || tv[v].is_on_bbox) continue;
            std::vector<int> ring;
            vs.oneRing((int)v, ring);
            double E;
            std::array<double, 3> J, X0;
            std::array<double, 9> H;
            bool ok = vs.NewtonsUpdate(ring, (int)v, E, J, H, X0);
            double ne = vs.getNewEnergy(ring);
            double rec[15] = {ok ? 1.0 : 0.0, E};
            std::memcpy(rec + 2, J.data(), 24);
            std::memcpy(rec + 5, H.data(), 72);
            rec[14] = ne;
            fwrite(rec, 8, 15, o);
            done++;
        }
        double T[12];
        for (int j = 0; j < 4; ++j) std::memcpy(T + 3 * j, tv[tets[0][j]].posf.data(), 24);
        double e0 = LocalOperations::comformalAMIPSEnergy_new(T);
        fwrite(&e0, 8, 1, o);
        fclose(o);
        printf("{\"counter\": %lld, \"suc_counter\": %lld, \"max_energy\": %.17g}\n",
               (long long)counters[0], (long long)counters[1], vs.getMaxEnergy());
    } catch (const std::exception& e) {
        fprintf(stderr, "shim_driver: %s\n", e.what());
        return 1;
    }
    return 0;
}
