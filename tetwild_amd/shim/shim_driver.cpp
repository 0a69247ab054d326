// shim_driver -- runs VertexSmoother / LocalOperations (the MI355X mirrors)
// on a mesh file and writes the result, so that a caller written against the
// reference's class surface can be exercised end to end.
//
//   shim_driver <mesh.bin> <out.bin> <mode>
//   mode: seq       smooth(): the reference's vertex-order (Gauss-Seidel) results,
//                   one launch per dependency level
//         serial    smoothSingleSerial() + smoothSurfaceSerial(): the same
//                   results, one launch per vertex
//         batch     smoothSingleBatched() + smoothSurfaceBatched()
//         collapse  EdgeCollapser::checkCollapses over every directed edge
//                   (v1, v2 ascending); collapse_soft: is_soft, soft_energy 40
//         remove    EdgeRemover::checkRemovals over every interior edge (closed
//                   ring) of 3..5 tets (edges ascending, rings = ascending tet ids);
//                   out.bin: per edge int32 kind, int32 k, k x (int32[4] tet,
//                   double quality); then int32 equal_buget
//         energy    getAvgMaxEnergy / getMaxEnergy / getSecondMaxEnergy /
//                   getFilterEnergy with tets t % 7 == 3 marked removed; prints
//                   them as JSON (out.bin unused)
//
// mesh.bin (little-endian): int64 nv, nt, nsf, nbf, is_mesh_closed;
// double eps, eps_2, sampling_dist; double posf[nv][3]; uint8 flags[nv]
// (1 surface, 2 bbox, 4 boundary, 8 locked, 16 unrounded); int32 tets[nt][4];
// int32 is_surface_fs[nt][4]; double tet_quality[nt];
// double sf_tri[nsf][9]; double b_tri[nbf][9]; then, for the unrounded
// vertices in vertex order, their exact positions double pos[][3].  The
// exact positions here are doubles, so the driver's ExactPositions orients
// them with the library's exact predicate (twg_tet_flip: positive or not).
// When the mesh has unrounded vertices, seq / serial / batch also write
// <out.bin>.rounded: uint8 is_rounded[nv] after the run.
// out.bin: double posf[nv][3]; double slim_energy[nt]; int64 counter,
// suc_counter; plus, for the first up-to-8 interior vertices, NewtonsUpdate
// (E, J[3], H[9]) and getNewEnergy at the final positions, and the static
// comformalAMIPSEnergy_new of tet 0.
// collapse out.bin: int32 codes[n]; int32 counts[n] (new qualities per
// candidate); double new_q[sum counts]; uint8 flags[nv] after the checks.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "tetwild_shim.h"

using namespace tetwild_amd;

template <class T>
static bool rd(FILE* f, T* p, size_t n)
{
    return fread(p, sizeof(T), n, f) == n;
}

// Exact positions of the unrounded vertices (doubles in these files): an
// orientation is the exact sign of the determinant on the device; the
// mirrors only ask whether it is positive.
class DoubleExact : public ExactPositions {
public:
    DoubleExact(Device& dev, const std::vector<TetVertex>& tv, std::vector<Point_3f> pos)
        : dev_(dev), tv_(tv), pos_(std::move(pos))
    {
    }
    int orientation(const std::array<int, 4>& t, int moved, const Point_3f& at) override
    {
        double xyz[12];
        for (int j = 0; j < 4; ++j) {
            const int v = t[j];
            const Point_3f& p = v == moved ? at : (tv_[v].is_rounded ? tv_[v].posf : pos_[v]);
            xyz[j] = p[0];
            xyz[4 + j] = p[1];
            xyz[8 + j] = p[2];
        }
        const int32_t ids[4] = {0, 1, 2, 3};
        char* b = (char*)dev_.buffer(5, 256);
        dev_.h2d(b, xyz, 96);
        dev_.h2d(b + 96, ids, 16);
        uint8_t* fl = (uint8_t*)(b + 128);
        if (twg_tet_flip(dev_.ctx(), (const int32_t*)(b + 96), 1, (double*)b, (double*)b + 4,
                         (double*)b + 8, 4, fl) != TWG_OK)
            throw std::runtime_error(std::string("twg_tet_flip: ") + twg_last_error());
        uint8_t f = 1;
        dev_.d2h(&f, fl, 1);
        calls++;
        return f ? -1 : 1;
    }
    void rounded(int v) override { pos_[v] = tv_[v].posf; }
    long calls = 0;

private:
    Device& dev_;
    const std::vector<TetVertex>& tv_;
    std::vector<Point_3f> pos_;
};

int main(int argc, char** argv)
{
    if (argc != 4) {
        fprintf(stderr, "usage: shim_driver mesh.bin out.bin seq|serial|batch|collapse|collapse_soft|remove|energy\n");
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
    std::vector<Point_3f> exact_pos(nv);
    int64_t n_unrounded = 0;
    for (int64_t v = 0; v < nv; ++v) {
        tv[v].posf = {{pos[3 * v], pos[3 * v + 1], pos[3 * v + 2]}};
        tv[v].is_on_surface = flags[v] & 1;
        tv[v].is_on_bbox = flags[v] & 2;
        tv[v].is_on_boundary = flags[v] & 4;
        tv[v].is_locked = flags[v] & 8;
        tv[v].is_rounded = !(flags[v] & 16);
        exact_pos[v] = tv[v].posf;
        if (!tv[v].is_rounded) {
            if (!rd(f, exact_pos[v].data(), 3)) return 2;
            n_unrounded++;
        }
    }
    fclose(f);
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
        DoubleExact exact(dev, tv, exact_pos);
        State st;
        st.eps = eps[0];
        st.eps_2 = eps[1];
        st.sampling_dist = eps[2];
        st.is_mesh_closed = hdr[4] != 0;
        std::string mode = argv[3];
        if (mode == "energy") {
            for (int64_t t = 0; t < nt; ++t) t_rm[t] = t % 7 == 3;
            LocalOperations lo(tv, tets, fs, v_rm, t_rm, tq, sf_tree, b_tree, st, dev);
            lo.exact = n_unrounded ? &exact : nullptr;
            double avg, mx;
            lo.getAvgMaxEnergy(avg, mx);
            bool clean = false;
            double fe = lo.getFilterEnergy(clean);
            printf("{\"avg\": %.17g, \"max\": %.17g, \"max2\": %.17g, \"second\": %.17g, "
                   "\"filter\": %.17g, \"clean\": %d}\n",
                   avg, mx, lo.getMaxEnergy(), lo.getSecondMaxEnergy(mx), fe, clean ? 1 : 0);
            return 0;
        }
        if (mode == "remove") {
            EdgeRemover er(tv, tets, fs, v_rm, t_rm, tq, sf_tree, b_tree, st, dev);
            er.exact = n_unrounded ? &exact : nullptr;
            std::vector<std::array<int, 2>> edges;
            for (const auto& t : tets)
                for (int a = 0; a < 4; ++a)
                    for (int b = a + 1; b < 4; ++b)
                        edges.push_back({{std::min(t[a], t[b]), std::max(t[a], t[b])}});
            std::sort(edges.begin(), edges.end());
            edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
            std::vector<std::array<int, 2>> cand;
            std::vector<std::vector<int>> rings;
            for (const auto& e : edges) {
                std::vector<int> r;
                for (int t : tv[e[0]].conn_tets)
                    if (tv[e[1]].conn_tets.count(t)) r.push_back(t);
                std::sort(r.begin(), r.end());
                if (r.size() < 3 || r.size() > 5) continue;
                // closed rings only (interior edges: every other vertex of the
                // ring appears in exactly two of its tets)
                std::map<int, int> cnt;
                for (int t : r)
                    for (int v : tets[t])
                        if (v != e[0] && v != e[1]) cnt[v]++;
                bool closed = true;
                for (const auto& kv : cnt) closed = closed && kv.second == 2;
                if (!closed) continue;
                cand.push_back(e);
                rings.push_back(r);
            }
            std::vector<EdgeRemover::Removal> res;
            er.checkRemovals(cand, rings, res);
            FILE* o = fopen(argv[2], "wb");
            if (!o) return 2;
            int n_ok = 0;
            for (const auto& r : res) {
                int32_t hdr2[2] = {r.kind, (int32_t)r.new_tets.size()};
                fwrite(hdr2, 4, 2, o);
                for (size_t k = 0; k < r.new_tets.size(); ++k) {
                    fwrite(r.new_tets[k].data(), 4, 4, o);
                    fwrite(&r.new_qs[k].slim_energy, 8, 1, o);
                }
                n_ok += r.kind != 0;
            }
            int32_t eb = er.equal_buget;
            fwrite(&eb, 4, 1, o);
            fclose(o);
            printf("{\"edges\": %zu, \"removable\": %d}\n", cand.size(), n_ok);
            return 0;
        }
        if (mode == "collapse" || mode == "collapse_soft") {
            EdgeCollapser ec(tv, tets, fs, v_rm, t_rm, tq, sf_tree, b_tree, st, dev);
            ec.exact = n_unrounded ? &exact : nullptr;
            if (mode == "collapse_soft") {
                ec.is_soft = true;
                ec.soft_energy = 40;
            }
            std::vector<std::array<int, 2>> edges;
            for (const auto& t : tets)
                for (int a = 0; a < 4; ++a)
                    for (int b = 0; b < 4; ++b)
                        if (a != b) edges.push_back({{t[a], t[b]}});
            std::sort(edges.begin(), edges.end());
            edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
            std::vector<int> codes;
            std::vector<std::vector<TetQuality>> qs;
            ec.checkCollapses(edges, codes, &qs);
            FILE* o = fopen(argv[2], "wb");
            if (!o) return 2;
            std::vector<int32_t> c32(codes.begin(), codes.end()), cnt;
            for (const auto& q : qs) cnt.push_back((int32_t)q.size());
            fwrite(c32.data(), 4, c32.size(), o);
            fwrite(cnt.data(), 4, cnt.size(), o);
            for (const auto& q : qs)
                for (const auto& t : q) fwrite(&t.slim_energy, 8, 1, o);
            for (int64_t v = 0; v < nv; ++v) {
                uint8_t fl = (tv[v].is_on_surface ? 1 : 0) | (tv[v].is_on_bbox ? 2 : 0) |
                             (tv[v].is_on_boundary ? 4 : 0) | (tv[v].is_locked ? 8 : 0) |
                             (tv[v].is_rounded ? 0 : 16);
                fwrite(&fl, 1, 1, o);
            }
            fclose(o);
            int ok = 0;
            for (int c : codes) ok += c == EdgeCollapser::SUCCESS || c == EdgeCollapser::ENVELOP_SUC;
            printf("{\"candidates\": %zu, \"collapsible\": %d, \"exact_calls\": %ld}\n",
                   edges.size(), ok, exact.calls);
            return 0;
        }
        VertexSmoother vs(tv, tets, fs, v_rm, t_rm, tq, sf_tree, b_tree, st, dev);
        vs.exact = n_unrounded ? &exact : nullptr;
        int64_t counters[2];
        int levels[2] = {0, 0};
        const auto t_smooth = std::chrono::steady_clock::now();
        if (mode == "seq") {
            // smooth(), recording the levels of its two loops
            vs.smoothSingle();
            levels[0] = vs.last_levels;
            if (st.eps >= 0) {
                vs.smoothSurface();
                levels[1] = vs.last_levels;
            }
        } else if (mode == "serial") {
            vs.smoothSingleSerial();
            if (st.eps >= 0) vs.smoothSurfaceSerial();
        } else {
            vs.smoothSingleBatched();
            if (st.eps >= 0) vs.smoothSurfaceBatched();
        }
        const double smooth_s =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t_smooth).count();
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
        if (n_unrounded) {
            FILE* r = fopen((std::string(argv[2]) + ".rounded").c_str(), "wb");
            if (!r) return 2;
            for (int64_t v = 0; v < nv; ++v) {
                uint8_t b = tv[v].is_rounded ? 1 : 0;
                fwrite(&b, 1, 1, r);
            }
            fclose(r);
        }
        printf("{\"counter\": %lld, \"suc_counter\": %lld, \"max_energy\": %.17g, "
               "\"exact_calls\": %ld, \"levels_single\": %d, \"levels_surface\": %d, "
               "\"smooth_seconds\": %.6f}\n",
               (long long)counters[0], (long long)counters[1], vs.getMaxEnergy(), exact.calls,
               levels[0], levels[1], smooth_s);
    } catch (const std::exception& e) {
        fprintf(stderr, "shim_driver: %s\n", e.what());
        return 1;
    }
    return 0;
}
