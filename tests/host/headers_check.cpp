// Host build of the device headers (twg_geom.hpp, twg_bvh.hpp) with the GPU
// qualifiers removed, checking the logic the kernels run:
//   1. SampleGen against the oracle's sampleTriangle (bit-exact samples and
//      counts), SampleGen::skip against stepping one sample at a time, and
//      the stored-record generator SampleGenC and the segmented layout
//      (store_seg, seg_row, SegFace) against both;
//   2. in_envelope -- global walker and LDS walker over the top levels, FP64
//      and FP32 boxes -- against the brute-force minimum over all facets, on
//      trees of 1..50k facets (the tree built by the library's own host code);
//   3. face_out_group's block split (G lanes emulated one after another, the
//      ballot as the OR of the lanes' step results) against face_out_sampled,
//      and that every sample of an inside triangle is tested exactly once;
//   4. collect_facets (one-lane group) against the facets whose boxes meet
//      the query box, in_envelope_list on its lists against brute force, and
//      in_envelope_masked with face_list_mask against in_envelope_list;
//   5. nearest_facet_group (one-lane group) against nearest_facet, bit for
//      bit, with queries above shared grid edges and vertices (ties);
//   6. point_tri_d2 (squared distance and closest point) against the
//      oracle's orc_point_tri_d2, bit for bit, with points in every region
//      of random, thin and degenerate triangles, on vertices and edges.
// Exit status 0 when every check passes.
#include "twg_bvh.hpp"

#include <cstdio>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "build_part.hpp"

extern "C" int64_t orc_sample_triangle(const double* v0, const double* v1, const double* v2,
                                       double d, double* ps, int64_t cap);
extern "C" double orc_point_tri_d2(const double* p, const double* a, const double* b, const double* c,
                                   double* cp);
using namespace twg;

static long g_bad = 0;
#define CHECK(c, ...)                  \
    do {                               \
        if (!(c)) {                    \
            if (g_bad < 10) printf(__VA_ARGS__); \
            g_bad++;                   \
        }                              \
    } while (0)

static void check_sampler(std::mt19937_64& rng)
{
    std::uniform_real_distribution<double> U(-1, 1);
    std::vector<double> ps(3 * 100000);
    long samples = 0, positions = 0, seg_faces = 0, seg_rows = 0;
    for (int it = 0; it < 6000; ++it) {
        double t[9];
        const double s = (it % 3 == 0) ? 0.01 : (it % 3 == 1 ? 0.003 : 0.02);
        for (int k = 0; k < 9; ++k) t[k] = U(rng) * s;
        if (it % 17 == 0)  // nearly collinear
            for (int k = 0; k < 3; ++k) t[6 + k] = t[k] + 0.5 * (t[3 + k] - t[k]) + (k == 2 ? 1e-4 : 0);
        if (it % 23 == 0)  // an edge of exactly 4 sampling distances
            for (int k = 0; k < 3; ++k) t[3 + k] = t[k] + (k == 0 ? 0.004 : 0.0);
        const double d = 0.001;
        const int64_t n = orc_sample_triangle(t, t + 3, t + 6, d, ps.data(), 100000);
        const V3<double> a = mk3(t[0], t[1], t[2]), b = mk3(t[3], t[4], t[5]), c = mk3(t[6], t[7], t[8]);
        std::vector<V3<double>> seq;
        for_each_sample(a, b, c, d, [&](V3<double> p) {
            seq.push_back(p);
            return (int64_t)seq.size() > n + 10;
        });
        CHECK((int64_t)seq.size() == n, "sampler: triangle %d gives %zu samples, oracle %lld\n", it,
              seq.size(), (long long)n);
        for (int64_t j = 0; j < n && j < (int64_t)seq.size(); ++j)
            CHECK(seq[j].x == ps[3 * j] && seq[j].y == ps[3 * j + 1] && seq[j].z == ps[3 * j + 2],
                  "sampler: triangle %d sample %lld differs\n", it, (long long)j);
        samples += n;
        SampleGen cnt;
        cnt.init(a, b, c, d);
        // the stored-record generator: the same sequence, bit for bit
        double rec[kFaceRec];
        for (double& v : rec) v = std::nan("");
        cnt.store(rec);
        {
            SampleGenC<const double*> gc;
            gc.load(rec, d);
            V3<double> p;
            int64_t j = 0;
            while (gc.next(p) && j <= n + 10) {
                CHECK(j < (int64_t)seq.size() && !memcmp(&p, &seq[j], sizeof p),
                      "stored sampler: triangle %d sample %lld differs\n", it, (long long)j);
                ++j;
            }
            CHECK(j == (int64_t)seq.size(), "stored sampler: triangle %d gives %lld samples\n", it,
                  (long long)j);
            SampleGenC<const double*> gcnt;
            gcnt.load(rec, d);
            CHECK(gcnt.skip(INT64_MAX) == (int64_t)seq.size(), "stored skip: count of triangle %d\n", it);
        }
        {
            // the segmented layout, filled as the smoothing kernel fills it:
            // every row on its own (seg_row), then the rows before the first
            // that ends and their samples
            double srec[kSegRec];
            int hdr[kSegHdr];
            for (double& v : srec) v = std::nan("");
            cnt.store_seg(srec, hdr, 1);  // row-table entry 0 unused
            std::vector<double> rt(kSegRow * (size_t)(hdr[kShRows] + 1), std::nan(""));
            for (int m = 1; m <= hdr[kShRows]; ++m) {
                V3<double> st = mk3(0.0, 0.0, 0.0);
                const int c = seg_row(srec, d, m, st);
                double* e = &rt[kSegRow * (size_t)m];
                e[0] = st.x;
                e[1] = st.y;
                e[2] = st.z;
                e[3] = c;
            }
            int rows = 0;
            int64_t sum = 0;
            while (rows < hdr[kShRows] && rt[kSegRow * (size_t)(rows + 1) + 3] > 0)
                sum += (int64_t)rt[kSegRow * (size_t)(++rows) + 3];
            hdr[kShRows] = rows;
            hdr[kShSum] = (int)sum;
            SegFace F;
            F.load(hdr);
            CHECK(F.total(sum) == (int64_t)seq.size(), "segments: triangle %d counts %lld of %zu\n", it,
                  (long long)F.total(sum), seq.size());
            int64_t j = 0;
            for (int s = 0; s < F.nseg(); ++s)
                for (int k = 0; k < F.count(rt.data(), s); ++k, ++j) {
                    const V3<double> p = F.sample(srec, rt.data(), s, k, d);
                    CHECK(j < (int64_t)seq.size() && !memcmp(&p, &seq[j], sizeof p),
                          "segments: triangle %d sample %lld (segment %d, %d) differs\n", it, (long long)j, s,
                          k);
                }
            CHECK(j == (int64_t)seq.size(), "segments: triangle %d walks %lld samples\n", it, (long long)j);
            for (size_t o = 0; o < seq.size(); o += 1 + seq.size() / 11) {
                int s, k;
                F.seek(rt.data(), (int64_t)o, s, k);
                const V3<double> p = F.sample(srec, rt.data(), s, k, d);
                CHECK(!memcmp(&p, &seq[o], sizeof p), "segments: triangle %d seek %zu\n", it, o);
            }
            seg_faces++;
            seg_rows += rows;
        }
        CHECK(cnt.skip(INT64_MAX) == (int64_t)seq.size(), "skip: count of triangle %d\n", it);
        const size_t step = it % 5 == 0 ? 1 : 1 + seq.size() / 7;
        for (size_t s0 = 0; s0 <= seq.size(); s0 += step) {
            SampleGen g;
            g.init(a, b, c, d);
            const int64_t got = g.skip((int64_t)s0);
            V3<double> p;
            const bool more = g.next(p);
            positions++;
            CHECK(got == (int64_t)s0, "skip: triangle %d skipped %lld of %zu\n", it, (long long)got, s0);
            if (s0 < seq.size())
                CHECK(more && !memcmp(&p, &seq[s0], sizeof p), "skip: triangle %d sample %zu\n", it, s0);
            else
                CHECK(!more, "skip: triangle %d past the end\n", it);
            SampleGenC<const double*> gc;
            gc.load(rec, d);
            const int64_t gotc = gc.skip((int64_t)s0);
            V3<double> pc;
            const bool morec = gc.next(pc);
            CHECK(gotc == (int64_t)s0, "stored skip: triangle %d skipped %lld of %zu\n", it,
                  (long long)gotc, s0);
            if (s0 < seq.size())
                CHECK(morec && !memcmp(&pc, &seq[s0], sizeof pc), "stored skip: triangle %d sample %zu\n",
                      it, s0);
            else
                CHECK(!morec, "stored skip: triangle %d past the end\n", it);
        }
    }
    printf("sampler: 6000 triangles, %ld samples, %ld skip positions; segments: %ld triangles, %ld rows\n",
           samples, positions, seg_faces, seg_rows);
}

struct Tree {
    std::vector<double> tri;
    std::vector<BvhNode64> n64;
    std::vector<BvhNode32> n32;
    BvhView view;
};

static void build_nodes(Tree& T);

static void build(Tree& T, int nf, std::mt19937_64& rng, double size)
{
    std::uniform_real_distribution<double> U(0, 1);
    T.tri.resize(9 * (size_t)nf);
    for (int f = 0; f < nf; ++f) {
        const double cx = U(rng), cy = U(rng), cz = 0.05 * sin(7 * cx) * cos(5 * cy);
        for (int k = 0; k < 3; ++k) {
            T.tri[9 * f + 3 * k] = cx + size * (U(rng) - 0.5);
            T.tri[9 * f + 3 * k + 1] = cy + size * (U(rng) - 0.5);
            T.tri[9 * f + 3 * k + 2] = cz + 0.25 * size * (U(rng) - 0.5);
        }
    }
    build_nodes(T);
}

// the tree over T.tri as the library builds it (facets in the given order)
static void build_nodes(Tree& T)
{
    const int nf = (int)(T.tri.size() / 9);
    const int64_t nnode = max_node(1, 0, nf) + 1;
    T.n64.assign(nnode, BvhNode64{});
    for (auto& N : T.n64) {
        for (int c = 0; c < 3; ++c) {
            N.lo[c] = INFINITY;
            N.hi[c] = -INFINITY;
        }
        N.leaf = -2;
    }
    int depth = 0;
    fill(T.n64, T.tri, 1, 0, nf, 0, depth);
    T.n32.resize(nnode);
    for (int64_t i = 0; i < nnode; ++i) {
        for (int c = 0; c < 3; ++c) {
            T.n32[i].lo[c] = round_down(T.n64[i].lo[c]);
            T.n32[i].hi[c] = round_up(T.n64[i].hi[c]);
        }
        T.n32[i].leaf = T.n64[i].leaf;
        T.n32[i].pad = 0;
    }
    T.view = {T.n64.data(), T.n32.data(), T.tri.data(), nf, nnode};
}

static void check_traversal(std::mt19937_64& rng)
{
    std::uniform_real_distribution<double> U(0, 1);
    std::vector<unsigned char> cache(kTreeCacheBytes);
    long queries = 0;
    for (int nf : {1, 2, 3, 7, 100, 1000, 3001, 20001}) {
        Tree T;
        build(T, nf, rng, 0.02);
        const double eps = 0.03, eps2 = eps * eps, Tt = eps2 * (1 + 1e-9), sqrtT = sqrt(Tt);
        for (int mode = 0; mode < 4; ++mode) {
            const bool f32 = mode & 1, lds = mode & 2;
            const BvhWalk W = lds ? (f32 ? cache_tree_top<true>(T.view, cache.data())
                                         : cache_tree_top<false>(T.view, cache.data()))
                                  : walk(T.view);
            for (int q = 0; q < 600; ++q) {
                const V3<double> p = mk3(U(rng), U(rng), 0.05 * sin(7 * U(rng)) + 0.01 * (U(rng) - 0.5));
                double best = INFINITY;
                for (int f = 0; f < nf; ++f) best = std::min(best, facet_d2(W, f, p));
                Stats st;
                int hint = (q % 3 == 0) ? (int)(rng() % nf) : -1;
                const bool got = f32 ? in_envelope<true, false>(W, p, eps2, Tt, sqrtT, hint, st)
                                     : in_envelope<false, false>(W, p, eps2, Tt, sqrtT, hint, st);
                CHECK(got == (best <= eps2), "traversal: %d facets, mode %d, query %d\n", nf, mode, q);
                queries++;
            }
        }
    }
    printf("traversal: %ld queries against brute force\n", queries);
}

template <int G>
static bool group_emulated(const BvhWalk& B, V3<double> a, V3<double> b, V3<double> c, double eps2,
                           double d, double Tt, double sqrtT, long& tested)
{
    const double l = fmax(len2(a - b), fmax(len2(b - c), len2(c - a)));
    if (!(sqrt(l) / d <= kMaxEdgeSamples)) return true;
    SampleGen cnt;
    cnt.init(a, b, c, d);
    const int64_t total = cnt.skip(INT64_MAX);
    const int64_t K = (total + G - 1) / G;
    std::vector<SampleGen> g(G);
    std::vector<int> hint(G, -1);
    Stats st;
    for (int r = 0; r < G; ++r) {
        g[r].init(a, b, c, d);
        g[r].skip((int64_t)r * K);
    }
    for (int64_t j = 0; j < K; ++j) {
        bool any = false;
        for (int r = 0; r < G; ++r) {
            V3<double> q = a;
            const bool have = (int64_t)r * K + j < total && g[r].next(q);
            tested += have;
            any |= have && !in_envelope<false, false>(B, q, eps2, Tt, sqrtT, hint[r], st);
        }
        if (any) return true;
    }
    return false;
}

static void check_group(std::mt19937_64& rng)
{
    std::uniform_real_distribution<double> U(0, 1);
    Tree T;
    build(T, 4000, rng, 0.03);
    const BvhWalk W = walk(T.view);
    const double eps = 0.01, eps2 = eps * eps, Tt = eps2 * (1 + 1e-9), sqrtT = sqrt(Tt);
    long outs = 0;
    for (int q = 0; q < 800; ++q) {
        const double cx = U(rng), cy = U(rng), s = 0.02 + 0.08 * U(rng);
        double t9[9];
        for (int k = 0; k < 3; ++k) {
            t9[3 * k] = cx + s * (U(rng) - 0.5);
            t9[3 * k + 1] = cy + s * (U(rng) - 0.5);
            t9[3 * k + 2] = 0.05 * sin(7 * cx) * cos(5 * cy) + 0.02 * (U(rng) - 0.5);
        }
        Stats st;
        const double d = eps * 0.5;
        const bool ref = face_out_sampled<false, false>(W, t9, eps2, d, Tt, sqrtT, st);
        int64_t total = 0;
        for_each_sample(ld3(t9), ld3(t9 + 3), ld3(t9 + 6), d, [&](V3<double>) {
            total++;
            return false;
        });
        long t64 = 0, t8 = 0, t2 = 0;
        const bool g64 = group_emulated<64>(W, ld3(t9), ld3(t9 + 3), ld3(t9 + 6), eps2, d, Tt, sqrtT, t64);
        const bool g8 = group_emulated<8>(W, ld3(t9), ld3(t9 + 3), ld3(t9 + 6), eps2, d, Tt, sqrtT, t8);
        const bool g2 = group_emulated<2>(W, ld3(t9), ld3(t9 + 3), ld3(t9 + 6), eps2, d, Tt, sqrtT, t2);
        CHECK(g64 == ref && g8 == ref && g2 == ref, "group: triangle %d decision\n", q);
        if (!ref) CHECK(t64 == total && t8 == total && t2 == total, "group: triangle %d coverage\n", q);
        outs += ref;
    }
    printf("group split: 800 triangles, %ld out\n", outs);
}

static void check_lists(std::mt19937_64& rng)
{
    std::uniform_real_distribution<double> U(0, 1);
    std::vector<unsigned char> cache(kTreeCacheBytes);
    long lists = 0, points = 0, overflow = 0, masked = 0, listed_faces = 0;
    for (int nf : {2, 3, 7, 100, 1000, 3001, 20001}) {
        Tree T;
        build(T, nf, rng, 0.02);
        const double eps = 0.01, eps2 = eps * eps, Tt = eps2 * (1 + 1e-9), sqrtT = sqrt(Tt);
        for (int mode = 0; mode < 3; ++mode) {
            const BvhWalk W = mode == 0 ? walk(T.view)
                                        : (mode == 1 ? cache_tree_top<false>(T.view, cache.data())
                                                     : cache_tree_top<true>(T.view, cache.data()));
            for (int q = 0; q < 100; ++q) {
                const int cap = q % 5 == 0 ? 16 : 4096;
                std::vector<BvhNode32> lst(cap);
                std::vector<uint32_t> fr(2 * cap);
                const double cx = U(rng), cy = U(rng), s = (q % 3 == 0) ? 0.3 : 0.04 * U(rng);
                double lo[3] = {cx, cy, 0.05 * sin(7 * cx) * cos(5 * cy) - 0.01};
                double hi[3] = {cx + s * U(rng), cy + s * U(rng), lo[2] + 0.02 + 0.03 * U(rng)};
                const double pad = 1.0001 * sqrtT + 1e-9;
                for (int k = 0; k < 3; ++k) {
                    lo[k] -= pad;
                    hi[k] += pad;
                }
                const int n = collect_facets<1>(W, lo, hi, lst.data(), fr.data(), cap, 0, 0);
                // exact: FP64 leaf boxes meeting the box; loose: FP32 leaf boxes
                std::vector<char> in_list(nf, 0), exact(nf, 0), loose(nf, 0);
                for (int64_t k = 1; k < T.view.nnode; ++k) {
                    const int f = T.n64[k].leaf;
                    if (f < 0) continue;
                    exact[f] = T.n64[k].lo[0] <= hi[0] && T.n64[k].lo[1] <= hi[1] && T.n64[k].lo[2] <= hi[2] &&
                               T.n64[k].hi[0] >= lo[0] && T.n64[k].hi[1] >= lo[1] && T.n64[k].hi[2] >= lo[2];
                    loose[f] = T.n32[k].lo[0] <= hi[0] + 1e-6 && T.n32[k].lo[1] <= hi[1] + 1e-6 &&
                               T.n32[k].lo[2] <= hi[2] + 1e-6 && T.n32[k].hi[0] >= lo[0] - 1e-6 &&
                               T.n32[k].hi[1] >= lo[1] - 1e-6 && T.n32[k].hi[2] >= lo[2] - 1e-6;
                }
                if (n < 0) {
                    // only a small capacity overflows (a work list can, with no facet in the box)
                    CHECK(cap < 4096, "lists: %d facets, mode %d, query %d: -1\n", nf, mode, q);
                    overflow++;
                    continue;
                }
                bool ok = true;
                for (int j = 0; j < n; ++j) {
                    const int f = lst[j].leaf;
                    ok = ok && f >= 0 && f < nf && !in_list[f] && loose[f];
                    if (f >= 0 && f < nf) in_list[f] = 1;
                }
                for (int f = 0; f < nf; ++f) ok = ok && (!exact[f] || in_list[f]);
                CHECK(ok, "lists: %d facets, mode %d, query %d: list differs from the box query\n", nf, mode, q);
                lists++;
                // half the lists record edge neighbours for in_envelope_list's
                // second try (any list positions: a facet within eps decides
                // "in" whichever facet it is, so arbitrary ones must not change
                // a decision)
                if (q % 2 == 1 && n < 256)
                    for (int j = 0; j < n; ++j) {
                        uint32_t pk = 0xff000000u;
                        for (int k = 0; k < 3; ++k) {
                            const uint32_t v = rng() % 4 == 0 ? 0xffu : (uint32_t)(rng() % n);
                            pk |= v << (8 * k);
                        }
                        lst[j].pad = (int32_t)pk;
                    }
                // points of the unpadded box against brute force
                std::vector<double> tri(9 * (size_t)std::max(n, 1));
                for (int j = 0; j < n; ++j)
                    for (int k = 0; k < 9; ++k) tri[9 * j + k] = T.tri[9 * (size_t)lst[j].leaf + k];
                int hl = -1;
                for (int r = 0; r < 16; ++r) {
                    V3<double> p;
                    double* pc[3] = {&p.x, &p.y, &p.z};
                    for (int k = 0; k < 3; ++k) *pc[k] = lo[k] + pad + (hi[k] - lo[k] - 2 * pad) * U(rng);
                    double best = INFINITY;
                    for (int f = 0; f < nf; ++f) best = std::min(best, facet_d2(W, f, p));
                    const bool got = q % 2 == 1
                                         ? in_envelope_list<true>(lst.data(), tri.data(), n, p, eps2, sqrtT, hl)
                                         : in_envelope_list(lst.data(), tri.data(), n, p, eps2, sqrtT, hl);
                    CHECK(got == (best <= eps2), "lists: %d facets, mode %d, query %d, point %d\n", nf, mode, q, r);
                    if (r % 7 == 0) hl = -1;
                    points++;
                }
                // a face inside the box, its samples tested through the face's
                // list mask: the decision and the facet left in the hint are
                // in_envelope_list's (the mask leaves out only facets farther
                // than sqrt(T) from the face)
                if (n >= 0 && n <= 128) {
                    V3<double> fv[3];
                    for (int j = 0; j < 3; ++j) {
                        double* pc[3] = {&fv[j].x, &fv[j].y, &fv[j].z};
                        for (int k = 0; k < 3; ++k) *pc[k] = lo[k] + pad + (hi[k] - lo[k] - 2 * pad) * U(rng);
                    }
                    uint64_t mk[2];
                    face_list_mask(lst.data(), n, fv[0], fv[1], fv[2], sqrtT, mk);
                    int hm = -1, hu = -1;
                    const bool nbr = q % 2 == 1;  // these lists carry neighbour positions
                    for_each_sample(fv[0], fv[1], fv[2], 0.004, [&](V3<double> p) {
                        // (the queries without neighbours take the four-at-a-time box reads)
                        const bool gm =
                            nbr ? in_envelope_masked<true>(lst.data(), tri.data(), n, mk, p, eps2, sqrtT, hm)
                                : in_envelope_masked<false, true>(lst.data(), tri.data(), n, mk, p, eps2, sqrtT,
                                                                  hm);
                        const bool gu =
                            nbr ? in_envelope_list<true>(lst.data(), tri.data(), n, p, eps2, sqrtT, hu)
                                : in_envelope_list(lst.data(), tri.data(), n, p, eps2, sqrtT, hu);
                        double best = INFINITY;
                        for (int f = 0; f < nf; ++f) best = std::min(best, facet_d2(W, f, p));
                        CHECK(gm == gu && hm == hu && gm == (best <= eps2),
                              "masked lists: %d facets, mode %d, query %d\n", nf, mode, q);
                        masked++;
                        return false;
                    });
                    // the whole decision of face_out_listed (masked scans)
                    // against the per-sample tree walk
                    double t9[9];
                    for (int j = 0; j < 3; ++j) {
                        t9[3 * j] = fv[j].x;
                        t9[3 * j + 1] = fv[j].y;
                        t9[3 * j + 2] = fv[j].z;
                    }
                    Stats st;
                    const bool lo_ = face_out_listed<false>(W, t9, eps2, 0.004, Tt, sqrtT, lst.data(),
                                                           tri.data(), n);
                    const bool ref =
                        face_out_sampled<false, false>(walk(T.view), t9, eps2, 0.004, Tt, sqrtT, st);
                    CHECK(lo_ == ref, "face_out_listed: %d facets, mode %d, query %d\n", nf, mode, q);
                    listed_faces++;
                }
            }
        }
    }
    printf("facet lists: %ld lists (%ld overflowed), %ld points against brute force, %ld face samples "
           "through face masks, %ld face_out_listed decisions\n", lists, overflow, points, masked,
           listed_faces);
}

// a grid of n x n squares over [0, 1]^2, each split into two triangles, on a
// ridge z = ridge * |x - 0.5| (0: a plane): queries above shared edges and
// vertices are equidistant from several facets
static void build_grid(Tree& T, int n, double ridge)
{
    T.tri.clear();
    auto z = [&](double x) { return ridge * fabs(x - 0.5); };
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            const double x0 = (double)i / n, x1 = (double)(i + 1) / n, y0 = (double)j / n, y1 = (double)(j + 1) / n;
            const double a[3] = {x0, y0, z(x0)}, b[3] = {x1, y0, z(x1)}, c[3] = {x1, y1, z(x1)}, d[3] = {x0, y1, z(x0)};
            for (const double* v : {a, b, c, a, c, d}) T.tri.insert(T.tri.end(), v, v + 3);
        }
    build_nodes(T);
}

static void check_nearest(std::mt19937_64& rng)
{
    std::uniform_real_distribution<double> U(0, 1);
    std::vector<unsigned char> cache(kTreeCacheBytes);
    long queries = 0, fallback = 0, off_hint = 0;
    for (int shape = 0; shape < 6; ++shape) {
        Tree T;
        if (shape < 3) build(T, shape == 0 ? 2 : (shape == 1 ? 700 : 6000), rng, 0.03);
        else build_grid(T, shape == 3 ? 9 : 24, shape == 5 ? 0.4 : 0.0);
        const int nf = (int)(T.tri.size() / 9);
        const int grid = shape == 4 ? 24 : (shape == 5 ? 24 : 9);
        for (int mode = 0; mode < 2; ++mode) {
            const BvhWalk W = mode == 0 ? walk(T.view) : cache_tree_top<false>(T.view, cache.data());
            for (int q = 0; q < 1500; ++q) {
                V3<double> p;
                if (shape >= 3 && q % 3 != 2) {
                    // above a grid vertex or on a grid line
                    const double x = (double)(rng() % (grid + 1)) / grid;
                    const double y = q % 3 == 0 ? (double)(rng() % (grid + 1)) / grid : U(rng);
                    p = mk3(x, y, (shape == 5 ? 0.4 * fabs(x - 0.5) : 0.0) + 0.05 * U(rng));
                } else {
                    p = mk3(1.4 * U(rng) - 0.2, 1.4 * U(rng) - 0.2, 0.2 * (U(rng) - 0.5));
                }
                if (q % 97 == 0) p.z += 50.0;  // far: the list overflows a small cap
                const int cap = q % 4 == 0 ? 8 : 256;
                std::vector<BvhNode32> lst(cap);
                std::vector<uint32_t> fr(2 * cap);
                V3<double> np0, np1;
                double sq0, sq1;
                const int f0 = nearest_facet(W, p, np0, sq0);
                const int f1 = nearest_facet_group<1>(W, p, np1, sq1, lst.data(), fr.data(), cap, 0, 0);
                CHECK(f0 == f1 && !memcmp(&np0, &np1, sizeof np0) && !memcmp(&sq0, &sq1, sizeof sq0),
                      "nearest: shape %d mode %d query %d: facet %d / %d\n", shape, mode, q, f0, f1);
                off_hint += f0 != nearest_hint(W, p);
                queries++;
            }
            (void)nf;
        }
    }
    (void)fallback;
    printf("nearest: %ld queries (%ld not at the hint facet) against the serial walk\n", queries, off_hint);
}

static void check_point_tri(std::mt19937_64& rng)
{
    std::uniform_real_distribution<double> U(-1, 1);
    long n = 0, edge_pts = 0;
    for (int it = 0; it < 400000; ++it) {
        double t[9], q[3];
        for (double& v : t) v = U(rng);
        const int kind = it % 10;
        if (kind == 7)  // thin: the third vertex near the first edge
            for (int k = 0; k < 3; ++k) t[6 + k] = t[k] + 0.37 * (t[3 + k] - t[k]) + 1e-7 * U(rng);
        if (kind == 8)  // collinear or a repeated vertex (the degenerate branch)
            for (int k = 0; k < 3; ++k) t[6 + k] = (it % 20 < 10) ? t[k] + 2.0 * (t[3 + k] - t[k]) : t[3 + k];
        const double spread = kind == 9 ? 0.01 : (kind < 3 ? 3.0 : 1.0);
        for (double& v : q) v = spread * U(rng);
        if (kind == 6) {  // on a vertex, or on an edge
            const int a = it % 3, b = (it + 1) % 3;
            const double w = (it % 6 < 3) ? 0.0 : 0.5 * (1 + U(rng));
            for (int k = 0; k < 3; ++k) q[k] = t[3 * a + k] + w * (t[3 * b + k] - t[3 * a + k]);
            edge_pts++;
        }
        double cpo[3];
        const double d2o = orc_point_tri_d2(q, t, t + 3, t + 6, cpo);
        V3<double> cp;
        const double d2 = point_tri_d2<double, true>(mk3(q[0], q[1], q[2]), mk3(t[0], t[1], t[2]),
                                                     mk3(t[3], t[4], t[5]), mk3(t[6], t[7], t[8]), cp);
        V3<double> cp2;
        const double d2b = point_tri_d2<double, false>(mk3(q[0], q[1], q[2]), mk3(t[0], t[1], t[2]),
                                                       mk3(t[3], t[4], t[5]), mk3(t[6], t[7], t[8]), cp2);
        CHECK(!memcmp(&d2, &d2o, 8) && !memcmp(&d2b, &d2o, 8), "point_tri_d2: case %d: %.17g vs %.17g\n", it,
              d2, d2o);
        CHECK(cp.x == cpo[0] && cp.y == cpo[1] && cp.z == cpo[2], "point_tri_d2: case %d closest point\n", it);
        n++;
    }
    printf("point_tri_d2: %ld cases (%ld on vertices and edges) vs orc_point_tri_d2\n", n, edge_pts);
}

int main()
{
    std::mt19937_64 rng(20240601);
    check_sampler(rng);
    check_traversal(rng);
    check_group(rng);
    check_lists(rng);
    check_nearest(rng);
    check_point_tri(rng);
    printf("%ld failures\n", g_bad);
    return g_bad != 0;
}
