// tetwild_shim.cpp -- host side of the LocalOperations / VertexSmoother
// mirrors.  Control flow follows src/tetwild/LocalOperations.cpp and
// src/tetwild/VertexSmoother.cpp; all geometry runs in libtetwild_amd.so.
#include "tetwild_shim.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>

namespace tetwild_amd {

namespace {
void check(int status, const char* what)
{
    if (status != TWG_OK)
        throw std::runtime_error(std::string(what) + ": " + twg_last_error());
}

void replace_id(std::array<int, 4>& t, int from, int to)
{
    for (int& v : t)
        if (v == from) {
            v = to;
            return;
        }
}

bool contains(const std::array<int, 4>& t, int v) { return std::find(t.begin(), t.end(), v) != t.end(); }
}  // namespace

// ---- Device --------------------------------------------------------------
Device::Device(int device) { check(twg_ctx_create(device, &ctx_), "twg_ctx_create"); }

Device::~Device()
{
    for (int i = 0; i < 8; ++i)
        if (buf_[i]) twg_free(ctx_, buf_[i]);
    twg_ctx_destroy(ctx_);
}

void* Device::buffer(int i, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    if (bytes > cap_[i]) {
        if (buf_[i]) check(twg_free(ctx_, buf_[i]), "twg_free");
        size_t want = std::max<size_t>(bytes + bytes / 2, 4096);
        check(twg_malloc(ctx_, want, &buf_[i]), "twg_malloc");
        cap_[i] = want;
    }
    return buf_[i];
}

void Device::h2d(void* dst, const void* src, size_t bytes)
{
    if (bytes) check(twg_memcpy_h2d(ctx_, dst, src, bytes), "twg_memcpy_h2d");
}

void Device::d2h(void* dst, const void* src, size_t bytes)
{
    if (bytes) check(twg_memcpy_d2h(ctx_, dst, src, bytes), "twg_memcpy_d2h");
}

void Device::sync() { check(twg_ctx_sync(ctx_), "twg_ctx_sync"); }

Device& default_device()
{
    static Device d(0);
    return d;
}

// ---- Envelope ------------------------------------------------------------
Envelope::Envelope(Device& dev, const std::vector<Point_3f>& V,
                   const std::vector<std::array<int, 3>>& F)
    : dev_(dev)
{
    check(twg_bvh_build(dev.ctx(), reinterpret_cast<const double*>(V.data()), (int64_t)V.size(),
                        reinterpret_cast<const int32_t*>(F.data()), (int64_t)F.size(), &bvh_),
          "twg_bvh_build");
}

Envelope::~Envelope() { twg_bvh_destroy(bvh_); }

double Envelope::squared_distance(const Point_3f& p, Point_3f* nearest) const
{
    // device buffers 6/7 are reserved for the envelope helpers
    double* dp = (double*)dev_.buffer(6, 64);
    double* dout = (double*)dev_.buffer(7, 64);
    dev_.h2d(dp, p.data(), 24);
    check(twg_nearest_points(dev_.ctx(), bvh_, dp, 1, dout, dout + 3, (int32_t*)(dout + 4)),
          "twg_nearest_points");
    double h[5];
    dev_.d2h(h, dout, 40);
    if (nearest) *nearest = {{h[0], h[1], h[2]}};
    return h[3];
}

// ---- LocalOperations -------------------------------------------------------
void LocalOperations::gather(const std::vector<std::array<int, 4>>& ts, Gathered& g, int pin,
                             int ov, const Point_3f* op)
{
    if (stamp_.size() < tet_vertices.size()) stamp_.assign(tet_vertices.size(), -1);
    g.tets.resize(4 * ts.size());
    g.global.clear();
    auto local = [&](int v) {
        if (stamp_[v] < 0) {
            stamp_[v] = (int)g.global.size();
            g.global.push_back(v);
        }
        return stamp_[v];
    };
    if (pin >= 0) local(pin);
    for (size_t i = 0; i < ts.size(); ++i)
        for (int j = 0; j < 4; ++j) g.tets[4 * i + j] = local(ts[i][j]);
    g.x.resize(g.global.size());
    g.y.resize(g.global.size());
    g.z.resize(g.global.size());
    for (size_t k = 0; k < g.global.size(); ++k) {
        const Point_3f& p = g.global[k] == ov && op ? *op : tet_vertices[g.global[k]].posf;
        g.x[k] = p[0];
        g.y[k] = p[1];
        g.z[k] = p[2];
        stamp_[g.global[k]] = -1;
    }
}

void LocalOperations::calTetQualities(const std::vector<std::array<int, 4>>& new_tets,
                                      std::vector<TetQuality>& tet_qs, bool)
{
    tet_qs.assign(new_tets.size(), TetQuality());
    if (new_tets.empty()) return;
    Gathered g;
    gather(new_tets, g);
    size_t n = new_tets.size(), nv = g.global.size();
    int32_t* dt = (int32_t*)dev.buffer(0, 16 * n);
    double* dx = (double*)dev.buffer(1, 8 * nv);
    double* dy = (double*)dev.buffer(2, 8 * nv);
    double* dz = (double*)dev.buffer(3, 8 * nv);
    double* de = (double*)dev.buffer(4, 8 * n);
    dev.h2d(dt, g.tets.data(), 16 * n);
    dev.h2d(dx, g.x.data(), 8 * nv);
    dev.h2d(dy, g.y.data(), 8 * nv);
    dev.h2d(dz, g.z.data(), 8 * nv);
    check(twg_tet_qualities(dev.ctx(), dt, (int64_t)n, dx, dy, dz, (int64_t)nv, TWG_FP64, 0.0, de,
                            nullptr),
          "twg_tet_qualities");
    std::vector<double> e(n);
    dev.d2h(e.data(), de, 8 * n);
    for (size_t i = 0; i < n; ++i) tet_qs[i].slim_energy = e[i];
}

// src/tetwild/LocalOperations.cpp:535-573
bool LocalOperations::isFlip(const std::vector<std::array<int, 4>>& new_tets)
{
    return isFlipAt(new_tets, -1, nullptr);
}

bool LocalOperations::hasUnroundedOther(const std::array<int, 4>& t, int v) const
{
    for (int u : t)
        if (u != v && !tet_vertices[u].is_rounded) return true;
    return false;
}

ExactPositions& LocalOperations::requireExact(const char* who) const
{
    if (!exact)
        throw std::logic_error(std::string(who) +
                               ": a tet has an unrounded vertex and no ExactPositions is set "
                               "(its orientation needs the exact rational positions)");
    return *exact;
}

bool LocalOperations::isFlipAt(const std::vector<std::array<int, 4>>& ts, int v, const Point_3f* at)
{
    std::vector<std::array<int, 4>> plain;
    std::vector<size_t> ex;
    plain.reserve(ts.size());
    for (size_t i = 0; i < ts.size(); ++i) {
        if (hasUnroundedOther(ts[i], v)) ex.push_back(i);
        else plain.push_back(ts[i]);
    }
    if (!plain.empty() && posfFlip(plain, v, at)) return true;
    if (!ex.empty()) {
        ExactPositions& X = requireExact("isFlip");
        const Point_3f here = v >= 0 ? (at ? *at : tet_vertices[v].posf) : Point_3f{{0, 0, 0}};
        for (size_t i : ex)
            if (X.orientation(ts[i], v, here) != 1) return true;
    }
    return false;
}

bool LocalOperations::posfFlip(const std::vector<std::array<int, 4>>& new_tets, int v,
                               const Point_3f* at)
{
    if (new_tets.empty()) return false;
    Gathered g;
    gather(new_tets, g, -1, v, at);
    size_t n = new_tets.size(), nv = g.global.size();
    int32_t* dt = (int32_t*)dev.buffer(0, 16 * n);
    double* dx = (double*)dev.buffer(1, 8 * nv);
    double* dy = (double*)dev.buffer(2, 8 * nv);
    double* dz = (double*)dev.buffer(3, 8 * nv);
    uint8_t* df = (uint8_t*)dev.buffer(4, n);
    dev.h2d(dt, g.tets.data(), 16 * n);
    dev.h2d(dx, g.x.data(), 8 * nv);
    dev.h2d(dy, g.y.data(), 8 * nv);
    dev.h2d(dz, g.z.data(), 8 * nv);
    check(twg_tet_flip(dev.ctx(), dt, (int64_t)n, dx, dy, dz, (int64_t)nv, df), "twg_tet_flip");
    std::vector<uint8_t> f(n);
    dev.d2h(f.data(), df, n);
    return std::any_of(f.begin(), f.end(), [](uint8_t b) { return b != 0; });
}

bool LocalOperations::isTetFlip(const std::array<int, 4>& t)
{
    return isFlipAt(std::vector<std::array<int, 4>>{t}, -1, nullptr);
}

void LocalOperations::qualitiesAt(const std::vector<std::array<int, 4>>& ts, int v,
                                  const Point_3f& at, std::vector<double>& q)
{
    q.assign(ts.size(), 0.0);
    if (ts.empty()) return;
    Gathered g;
    gather(ts, g, -1, v, &at);
    size_t n = ts.size(), nv = g.global.size();
    int32_t* dt = (int32_t*)dev.buffer(0, 16 * n);
    double* dx = (double*)dev.buffer(1, 8 * nv);
    double* dy = (double*)dev.buffer(2, 8 * nv);
    double* dz = (double*)dev.buffer(3, 8 * nv);
    double* de = (double*)dev.buffer(4, 8 * n);
    dev.h2d(dt, g.tets.data(), 16 * n);
    dev.h2d(dx, g.x.data(), 8 * nv);
    dev.h2d(dy, g.y.data(), 8 * nv);
    dev.h2d(dz, g.z.data(), 8 * nv);
    check(twg_tet_qualities(dev.ctx(), dt, (int64_t)n, dx, dy, dz, (int64_t)nv, TWG_FP64, 0.0, de,
                            nullptr),
          "twg_tet_qualities");
    dev.d2h(q.data(), de, 8 * n);
}

// src/tetwild/LocalOperations.cpp:575-603
void LocalOperations::getCheckQuality(const std::vector<TetQuality>& tet_qs, TetQuality& tq)
{
    double slim_sum = 0, slim_max = 0;
    for (const auto& q : tet_qs) {
        if (state.use_energy_max) {
            if (q.slim_energy > slim_max) slim_max = q.slim_energy;
        } else {
            slim_sum += q.slim_energy * q.volume;
        }
    }
    tq.slim_energy = state.use_energy_max ? slim_max : slim_sum;
}

void LocalOperations::getCheckQuality(const std::vector<int>& t_ids, TetQuality& tq)
{
    std::vector<TetQuality> qs;
    qs.reserve(t_ids.size());
    for (int t : t_ids) qs.push_back(tet_qualities[t]);
    getCheckQuality(qs, tq);
}

// One device pass for the energy reductions (src/tetwild/LocalOperations.cpp:
// 605-693): slim energies, removal flags, tets and vertex locks uploaded, the
// sums, maxima and the getFilterEnergy histogram computed by twg_energy_stats.
twg_energy_summary LocalOperations::energySummary()
{
    const size_t n = tet_qualities.size(), nv = tet_vertices.size();
    std::vector<double> e(n);
    std::vector<uint8_t> rm(n), lk(nv);
    std::vector<int32_t> ts(4 * n);
    for (size_t i = 0; i < n; ++i) {
        e[i] = tet_qualities[i].slim_energy;
        rm[i] = i < t_is_removed.size() && t_is_removed[i] ? 1 : 0;
        for (int j = 0; j < 4; ++j) ts[4 * i + j] = tets[i][j];
    }
    for (size_t v = 0; v < nv; ++v) lk[v] = tet_vertices[v].is_locked ? 1 : 0;
    double* de = (double*)dev.buffer(0, 8 * n + 8);
    uint8_t* drm = (uint8_t*)dev.buffer(1, n + nv + 16);
    uint8_t* dlk = drm + n;
    int32_t* dts = (int32_t*)dev.buffer(2, 16 * n + 16);
    dev.h2d(de, e.data(), 8 * n);
    dev.h2d(drm, rm.data(), n);
    dev.h2d(dlk, lk.data(), nv);
    dev.h2d(dts, ts.data(), 16 * n);
    twg_energy_summary out;
    check(twg_energy_stats(dev.ctx(), de, (int64_t)n, drm, dts, dlk, (int64_t)nv,
                           state.filter_energy_thres, &out),
          "twg_energy_stats");
    return out;
}

// src/tetwild/LocalOperations.cpp:605-622
void LocalOperations::getAvgMaxEnergy(double& avg_tq, double& max_tq)
{
    twg_energy_summary s = energySummary();
    max_tq = s.max;
    avg_tq = s.sum / (double)s.count;
    if (std::isinf(avg_tq)) avg_tq = State::MAX_ENERGY;
}

// src/tetwild/LocalOperations.cpp:624-635
double LocalOperations::getMaxEnergy() { return energySummary().max; }

// src/tetwild/LocalOperations.cpp:637-650 (the argument is unused there too)
double LocalOperations::getSecondMaxEnergy(double) { return energySummary().second_max; }

// src/tetwild/LocalOperations.cpp:652-693: the decision from the histogram
double LocalOperations::getFilterEnergy(bool& is_clean_up)
{
    twg_energy_summary s = energySummary();
    int64_t below[10], above[10];
    for (int i = 0; i < 10; ++i) {
        below[i] = above[i] = 0;
        for (int k = 0; k < 11; ++k) (k <= i ? below[i] : above[i]) += s.buckets[k];
    }
    if (below[0] >= above[0]) {
        is_clean_up = above[5] > 0;
        return 8;
    }
    if (below[8] <= above[8]) return 1e11;
    for (int i = 0; i < 8; ++i)
        if (below[i] < above[i] && below[i + 1] > above[i + 1])
            return state.filter_energy_thres - 1 + 5 * std::pow(10, i + 1);
    return 8;
}

void LocalOperations::areFacesOutEnvelop(const std::vector<Triangle_3f>& tris,
                                         std::vector<uint8_t>& out)
{
    out.assign(tris.size(), 0);
    if (tris.empty()) return;
    double* dt = (double*)dev.buffer(6, 72 * tris.size());
    uint8_t* dout = (uint8_t*)dev.buffer(7, tris.size());
    dev.h2d(dt, tris.data(), 72 * tris.size());
    check(twg_envelope_faces(dev.ctx(), geo_sf_tree.bvh(), dt, (int64_t)tris.size(), state.eps_2,
                             state.sampling_dist, TWG_FP64, dout, nullptr),
          "twg_envelope_faces");
    dev.d2h(out.data(), dout, tris.size());
}

void LocalOperations::arePointsOutEnvelop(const std::vector<Point_3f>& ps,
                                          std::vector<uint8_t>& out)
{
    out.assign(ps.size(), 0);
    if (ps.empty()) return;
    double* dp = (double*)dev.buffer(6, 24 * ps.size());
    uint8_t* dout = (uint8_t*)dev.buffer(7, ps.size());
    dev.h2d(dp, ps.data(), 24 * ps.size());
    check(twg_envelope_points(dev.ctx(), geo_sf_tree.bvh(), dp, (int64_t)ps.size(), state.eps_2,
                              TWG_FP64, dout, nullptr),
          "twg_envelope_points");
    dev.d2h(out.data(), dout, ps.size());
}

// src/tetwild/LocalOperations.cpp:967-975, 1046-1109
bool LocalOperations::isFaceOutEnvelop(const Triangle_3f& tri)
{
    std::vector<uint8_t> out;
    areFacesOutEnvelop({tri}, out);
    return out[0] != 0;
}

// src/tetwild/LocalOperations.cpp:1034-1044
bool LocalOperations::isPointOutEnvelop(const Point_3f& p)
{
    std::vector<uint8_t> out;
    arePointsOutEnvelop({p}, out);
    return out[0] != 0;
}

// src/tetwild/LocalOperations.cpp:1111-1121
bool LocalOperations::isPointOutBoundaryEnvelop(const Point_3f& p)
{
    double* dp = (double*)dev.buffer(6, 24);
    uint8_t* dout = (uint8_t*)dev.buffer(7, 1);
    dev.h2d(dp, p.data(), 24);
    check(twg_envelope_points(dev.ctx(), geo_b_tree.bvh(), dp, 1, state.eps_2, TWG_FP64, dout,
                              nullptr),
          "twg_envelope_points");
    uint8_t o = 0;
    dev.d2h(&o, dout, 1);
    return o != 0;
}

// src/tetwild/LocalOperations.cpp:1278-1287
bool LocalOperations::isIsolated(int v_id)
{
    for (int t : tet_vertices[v_id].conn_tets)
        for (int j = 0; j < 4; ++j)
            if (tets[t][j] != v_id && is_surface_fs[t][j] != State::NOT_SURFACE) return false;
    return true;
}

// src/tetwild/LocalOperations.cpp:1289-1303
bool LocalOperations::isBoundaryPoint(int v_id)
{
    if (state.is_mesh_closed) return false;
    std::unordered_set<int> n_v_ids;
    for (int t : tet_vertices[v_id].conn_tets)
        for (int j = 0; j < 4; ++j)
            if (tets[t][j] != v_id && tet_vertices[tets[t][j]].is_on_boundary)
                n_v_ids.insert(tets[t][j]);
    for (int n : n_v_ids)
        if (isEdgeOnBoundary(n, v_id)) return true;
    return false;
}

// src/tetwild/LocalOperations.cpp:931-965
bool LocalOperations::isEdgeOnBoundary(int v1_id, int v2_id)
{
    if (state.is_mesh_closed) return false;
    if (!tet_vertices[v1_id].is_on_boundary || !tet_vertices[v2_id].is_on_boundary) return false;
    int cnt = 0;
    for (int t : tet_vertices[v1_id].conn_tets) {
        std::array<int, 4> opp{};
        int ii = 0;
        for (int j = 0; j < 4; ++j) {
            if (tets[t][j] == v1_id || tets[t][j] == v2_id) continue;
            opp[ii++] = j;
        }
        if (ii == 2) {
            if (is_surface_fs[t][opp[0]] != State::NOT_SURFACE) cnt++;
            if (is_surface_fs[t][opp[1]] != State::NOT_SURFACE) cnt++;
            if (cnt > 2) return false;
        }
    }
    return cnt == 2;
}

// ---- static AMIPS functions (one tet, on the default device) --------------
namespace {
void amips_one(const double* T, int mode, double* E, double* G, double* H)
{
    Device& d = default_device();
    double xyz[12];
    for (int j = 0; j < 4; ++j) {
        xyz[j] = T[3 * j];
        xyz[4 + j] = T[3 * j + 1];
        xyz[8 + j] = T[3 * j + 2];
    }
    int32_t tet[4] = {0, 1, 2, 3};
    double* dx = (double*)d.buffer(0, 12 * 8);
    int32_t* dt = (int32_t*)d.buffer(1, 16);
    double* dout = (double*)d.buffer(2, 13 * 8);
    d.h2d(dx, xyz, 96);
    d.h2d(dt, tet, 16);
    check(twg_amips_eval(d.ctx(), dt, 1, dx, dx + 4, dx + 8, 4, mode, dout, dout + 1, dout + 4),
          "twg_amips_eval");
    double h[13];
    d.d2h(h, dout, 104);
    if (E) *E = h[0];
    if (G) std::memcpy(G, h + 1, 24);
    if (H) std::memcpy(H, h + 4, 72);
}
}  // namespace

double LocalOperations::comformalAMIPSEnergy_new(const double* T)
{
    double e;
    amips_one(T, TWG_AMIPS_E, &e, nullptr, nullptr);
    return e;
}

void LocalOperations::comformalAMIPSJacobian_new(const double* T, double* result_0)
{
    amips_one(T, TWG_AMIPS_EG, nullptr, result_0, nullptr);
}

void LocalOperations::comformalAMIPSHessian_new(const double* T, double* result_0)
{
    amips_one(T, TWG_AMIPS_EGH, nullptr, nullptr, result_0);
}

// ---- VertexSmoother --------------------------------------------------------
void VertexSmoother::oneRing(int v_id, std::vector<int>& t_ids) const
{
    t_ids.assign(tet_vertices[v_id].conn_tets.begin(), tet_vertices[v_id].conn_tets.end());
    std::sort(t_ids.begin(), t_ids.end());
}

void VertexSmoother::runBatch(const std::vector<Problem>& probs, std::vector<uint8_t>& status)
{
    size_t n = probs.size();
    status.assign(n, 0);
    if (n == 0) return;
    // gather every ring into one compact vertex set; centres first
    if (stamp_.size() < tet_vertices.size()) stamp_.assign(tet_vertices.size(), -1);
    std::vector<int> global;
    auto local = [&](int v) {
        if (stamp_[v] < 0) {
            stamp_[v] = (int)global.size();
            global.push_back(v);
        }
        return stamp_[v];
    };
    std::vector<int32_t> center(n), ring_off(n + 1), tri_off(n + 1);
    std::vector<uint8_t> kind(n);
    std::vector<std::vector<int>> rings(n);
    for (size_t p = 0; p < n; ++p) center[p] = local(probs[p].v);
    std::vector<int32_t> ring_tets, tri_ab;
    std::vector<double> old_q;
    ring_off[0] = 0;
    tri_off[0] = 0;
    for (size_t p = 0; p < n; ++p) {
        int v = probs[p].v;
        kind[p] = (uint8_t)probs[p].kind;
        oneRing(v, rings[p]);
        for (int t : rings[p]) {
            for (int j = 0; j < 4; ++j) ring_tets.push_back(local(tets[t][j]));
            old_q.push_back(tet_qualities.size() > (size_t)t ? tet_qualities[t].slim_energy : 0.0);
        }
        ring_off[p + 1] = (int32_t)(ring_tets.size() / 4);
        if (probs[p].kind == 1 || probs[p].kind == 2) {
            // one-ring surface faces, src/tetwild/VertexSmoother.cpp:316-327, 395-402
            std::vector<std::array<int, 3>> tri_ids;
            for (int t : rings[p])
                for (int j = 0; j < 4; ++j)
                    if (tets[t][j] != v && is_surface_fs[t][j] != State::NOT_SURFACE) {
                        std::array<int, 3> tri = {{tets[t][(j + 1) % 4], tets[t][(j + 2) % 4],
                                                   tets[t][(j + 3) % 4]}};
                        std::sort(tri.begin(), tri.end());
                        tri_ids.push_back(tri);
                    }
            std::sort(tri_ids.begin(), tri_ids.end());
            tri_ids.erase(std::unique(tri_ids.begin(), tri_ids.end()), tri_ids.end());
            for (const auto& tri : tri_ids) {
                int k = (int)(std::find(tri.begin(), tri.end(), v) - tri.begin());
                tri_ab.push_back(local(tri[(k + 1) % 3]));
                tri_ab.push_back(local(tri[(k + 2) % 3]));
            }
        }
        tri_off[p + 1] = (int32_t)(tri_ab.size() / 2);
    }
    size_t nv = global.size(), nr = ring_tets.size() / 4, nt = tri_ab.size() / 2;
    std::vector<double> x(nv), y(nv), z(nv);
    for (size_t k = 0; k < nv; ++k) {
        const Point_3f& p = tet_vertices[global[k]].posf;
        x[k] = p[0];
        y[k] = p[1];
        z[k] = p[2];
        stamp_[global[k]] = -1;
    }
    // device copies (buffer slots 0..5; 6/7 are the envelope helpers')
    char* b0 = (char*)dev.buffer(0, 24 * nv);
    double *dx = (double*)b0, *dy = dx + nv, *dz = dy + nv;
    int32_t* dring = (int32_t*)dev.buffer(1, 16 * nr + 16);
    char* b2 = (char*)dev.buffer(2, 4 * (n + (n + 1) + (n + 1)) + 2 * 4 * nt + 64);
    int32_t* dcenter = (int32_t*)b2;
    int32_t* droff = dcenter + n;
    int32_t* dtoff = droff + (n + 1);
    int32_t* dtri = dtoff + (n + 1);
    uint8_t* dkind = (uint8_t*)dev.buffer(3, n + n);
    uint8_t* dstatus = dkind + n;
    double* dq = (double*)dev.buffer(4, 16 * nr + 16);
    double* dnewq = dq + nr;
    double* dpout = (double*)dev.buffer(5, 24 * n);
    dev.h2d(dx, x.data(), 8 * nv);
    dev.h2d(dy, y.data(), 8 * nv);
    dev.h2d(dz, z.data(), 8 * nv);
    dev.h2d(dring, ring_tets.data(), 16 * nr);
    dev.h2d(dcenter, center.data(), 4 * n);
    dev.h2d(droff, ring_off.data(), 4 * (n + 1));
    dev.h2d(dtoff, tri_off.data(), 4 * (n + 1));
    dev.h2d(dtri, tri_ab.data(), 8 * nt);
    dev.h2d(dkind, kind.data(), n);
    dev.h2d(dq, old_q.data(), 8 * nr);
    twg_smooth_desc d;
    std::memset(&d, 0, sizeof d);
    d.x = dx;
    d.y = dy;
    d.z = dz;
    d.nv = (int64_t)nv;
    d.n = (int64_t)n;
    d.center = dcenter;
    d.ring_off = droff;
    d.ring_tets = dring;
    d.n_ring = (int64_t)nr;
    d.kind = dkind;
    d.old_q = dq;
    d.tri_off = dtoff;
    d.tri_ab = dtri;
    d.n_tri = (int64_t)nt;
    d.sf_tree = geo_sf_tree.bvh();
    d.b_tree = geo_b_tree.bvh();
    d.eps2 = state.eps_2;
    d.sampling_dist = state.sampling_dist > 0 ? state.sampling_dist : 1.0;
    d.p_out = dpout;
    d.status = dstatus;
    d.new_q = dnewq;
    check(twg_smooth_batch(dev.ctx(), &d, precision, 0, 0), "twg_smooth_batch");
    std::vector<double> pout(3 * n), newq(nr);
    dev.d2h(status.data(), dstatus, n);
    dev.d2h(pout.data(), dpout, 24 * n);
    dev.d2h(newq.data(), dnewq, 8 * nr);
    for (size_t p = 0; p < n; ++p) {
        if (status[p] == TWG_SMOOTH_BAD_INPUT)
            throw std::runtime_error("twg_smooth_batch: bad input for vertex " +
                                     std::to_string(probs[p].v));
        if (probs[p].kind == 3) continue;  // NewtonsMethod only: caller reads the position
        if (status[p] != TWG_SMOOTH_MOVED) continue;
        tet_vertices[probs[p].v].posf = {{pout[3 * p], pout[3 * p + 1], pout[3 * p + 2]}};
        if (probs[p].kind == 1 || probs[p].kind == 2) {
            int r0 = ring_off[p];
            for (size_t i = 0; i < rings[p].size(); ++i) {
                TetQuality q;
                q.slim_energy = newq[r0 + i];
                tet_qualities[rings[p][i]] = q;
            }
        }
    }
    last_pout_.assign(pout.begin(), pout.end());
}

// src/tetwild/VertexSmoother.cpp:465-542
bool VertexSmoother::NewtonsMethod(const std::vector<int>& t_ids,
                                   const std::vector<std::array<int, 4>>&, int v_id, Point_3f& p)
{
    (void)t_ids;  // the ring is the vertex's conn_tets (sorted), as oneRing() gathers it
    std::vector<int> ring;
    oneRing(v_id, ring);
    for (int t : ring)
        if (hasUnroundedOther(tets[t], v_id)) return newtonExact(v_id, ring, p);
    std::vector<uint8_t> st;
    runBatch({Problem{v_id, 3}}, st);
    p = {{last_pout_[0], last_pout_[1], last_pout_[2]}};
    return st[0] == TWG_SMOOTH_MOVED;
}

// src/tetwild/VertexSmoother.cpp:627-702
bool VertexSmoother::NewtonsUpdate(const std::vector<int>& t_ids, int v_id, double& energy,
                                   std::array<double, 3>& J, std::array<double, 9>& H,
                                   std::array<double, 3>& X0)
{
    std::vector<std::array<int, 4>> rot;
    for (int t : t_ids) {
        int start = 0;
        for (int j = 0; j < 4; ++j)
            if (tets[t][j] == v_id) {
                start = j;
                break;
            }
        rot.push_back({{tets[t][start], tets[t][(start + 1) % 4], tets[t][(start + 2) % 4],
                        tets[t][(start + 3) % 4]}});
    }
    Gathered g;
    gather(rot, g);
    size_t n = rot.size(), nv = g.global.size();
    int32_t* dt = (int32_t*)dev.buffer(0, 16 * n);
    double* dx = (double*)dev.buffer(1, 24 * nv);
    double* dout = (double*)dev.buffer(4, 104 * n);
    dev.h2d(dt, g.tets.data(), 16 * n);
    dev.h2d(dx, g.x.data(), 8 * nv);
    dev.h2d(dx + nv, g.y.data(), 8 * nv);
    dev.h2d(dx + 2 * nv, g.z.data(), 8 * nv);
    check(twg_amips_eval(dev.ctx(), dt, (int64_t)n, dx, dx + nv, dx + 2 * nv, (int64_t)nv,
                         TWG_AMIPS_EGH, dout, dout + n, dout + 4 * n),
          "twg_amips_eval");
    std::vector<double> h(13 * n);
    dev.d2h(h.data(), dout, 104 * n);
    energy = 0;
    J = {{0, 0, 0}};
    H = {{0, 0, 0, 0, 0, 0, 0, 0, 0}};
    X0 = tet_vertices[v_id].posf;
    for (size_t i = 0; i < n; ++i) {
        energy += h[i];
        for (int k = 0; k < 3; ++k) J[k] += h[n + 3 * i + k];
        for (int k = 0; k < 9; ++k) H[k] += h[4 * n + 9 * i + k];
    }
    if (std::isinf(energy)) energy = State::MAX_ENERGY;
    if (std::isnan(energy) || energy <= 0) return false;
    for (double v : J)
        if (!std::isfinite(v)) return false;
    for (double v : H)
        if (!std::isfinite(v)) return false;
    return true;
}

// src/tetwild/VertexSmoother.cpp:544-625 (scalar branch)
double VertexSmoother::getNewEnergy(const std::vector<int>& t_ids)
{
    std::vector<std::array<int, 4>> ts;
    for (int t : t_ids) ts.push_back(tets[t]);
    if (ts.empty()) return State::MAX_ENERGY;
    Gathered g;
    gather(ts, g);
    size_t n = ts.size(), nv = g.global.size();
    int32_t* dt = (int32_t*)dev.buffer(0, 16 * n);
    double* dx = (double*)dev.buffer(1, 24 * nv);
    double* de = (double*)dev.buffer(4, 8 * n);
    dev.h2d(dt, g.tets.data(), 16 * n);
    dev.h2d(dx, g.x.data(), 8 * nv);
    dev.h2d(dx + nv, g.y.data(), 8 * nv);
    dev.h2d(dx + 2 * nv, g.z.data(), 8 * nv);
    check(twg_amips_eval(dev.ctx(), dt, (int64_t)n, dx, dx + nv, dx + 2 * nv, (int64_t)nv,
                         TWG_AMIPS_E, de, nullptr, nullptr),
          "twg_amips_eval");
    std::vector<double> e(n);
    dev.d2h(e.data(), de, 8 * n);
    double s = 0;
    for (double v : e) s += v;
    if (std::isinf(s) || std::isnan(s) || s <= 0 || s > State::MAX_ENERGY) s = State::MAX_ENERGY;
    return s;
}

// smoothSingle / smoothSurface's try-to-round step
// (src/tetwild/VertexSmoother.cpp:152-161, 272-281): pos = Point_3(posf), kept
// when no ring tet flips
void VertexSmoother::tryRound(int v, const std::vector<std::array<int, 4>>& ring)
{
    TetVertex& tv = tet_vertices[v];
    if (tv.is_rounded) return;
    const Point_3f at = tv.posf;
    if (!isFlipAt(ring, v, &at)) {
        tv.is_rounded = true;
        if (exact) exact->rounded(v);
    }
}

// A moved vertex ends rounded: smoothSingle sets pos = p, posf = pf,
// is_rounded = true (:196-211); smoothSurface rounds it after the envelope
// test (:441-447), whose isFlip repeats the one the move already passed.
void VertexSmoother::afterMove(int v, int status)
{
    TetVertex& tv = tet_vertices[v];
    if (status == TWG_SMOOTH_MOVED && !tv.is_rounded) {
        tv.is_rounded = true;
        if (exact) exact->rounded(v);
    }
}

int VertexSmoother::smoothOne(int v, int kind)
{
    std::vector<int> ring;
    oneRing(v, ring);
    std::vector<std::array<int, 4>> rt;
    rt.reserve(ring.size());
    for (int t : ring) rt.push_back(tets[t]);
    tryRound(v, rt);
    bool needs_exact = false;
    for (const auto& t : rt) needs_exact = needs_exact || hasUnroundedOther(t, v);
    int s;
    if (needs_exact) {
        s = smoothExact(v, kind, ring);
    } else {
        std::vector<uint8_t> st;
        runBatch({Problem{v, kind}}, st);
        s = st[0];
    }
    afterMove(v, s);
    return s;
}

// NewtonsMethod (src/tetwild/VertexSmoother.cpp:465-542) for a ring whose flip
// tests partly need ExactPositions: one twg_newton_step launch per step (its
// device flip tests skip the exact tets), the halving loop on the host.
bool VertexSmoother::newtonExact(int v, const std::vector<int>& ring, Point_3f& p)
{
    ExactPositions& X = requireExact("NewtonsMethod");
    std::vector<std::array<int, 4>> rt;
    for (int t : ring) rt.push_back(tets[t]);
    const size_t m = rt.size();
    std::vector<uint8_t> skip(m);
    for (size_t j = 0; j < m; ++j) skip[j] = hasUnroundedOther(rt[j], v) ? 1 : 0;
    Gathered g;
    gather(rt, g, v);
    const size_t nv = g.global.size();
    char* b0 = (char*)dev.buffer(0, 24 * nv + 16 * m + m + sizeof(twg_newton_step_out) + 64);
    double* dx = (double*)b0;
    int32_t* dt = (int32_t*)(dx + 3 * nv);  // 24 nv is a multiple of 8: round up to 16
    dt = (int32_t*)(((uintptr_t)dt + 15) & ~(uintptr_t)15);
    uint8_t* dskip = (uint8_t*)(dt + 4 * m);
    twg_newton_step_out* dout =
        (twg_newton_step_out*)(((uintptr_t)(dskip + m) + 15) & ~(uintptr_t)15);
    dev.h2d(dx, g.x.data(), 8 * nv);
    dev.h2d(dx + nv, g.y.data(), 8 * nv);
    dev.h2d(dx + 2 * nv, g.z.data(), 8 * nv);
    dev.h2d(dt, g.tets.data(), 16 * m);
    dev.h2d(dskip, skip.data(), m);
    Point_3f pos = tet_vertices[v].posf;
    bool is_moved = false;
    twg_newton_step_out o;
    for (int step = 0; step < 15; ++step) {
        check(twg_newton_step(dev.ctx(), dx, dx + nv, dx + 2 * nv, (int64_t)nv, dt, (int32_t)m,
                              0, pos[0], pos[1], pos[2], dskip, dout),
              "twg_newton_step");
        dev.d2h(&o, dout, sizeof o);
        if (o.ok < 0) throw std::runtime_error("twg_newton_step: bad input");
        if (!o.ok) break;
        bool taken = false, tiny = false;
        for (int k = 0; k < 20 && !taken; ++k) {
            if (!o.finite[k] || o.flip[k]) continue;
            const Point_3f Xk = {{o.X[k][0], o.X[k][1], o.X[k][2]}};
            bool flip = false;
            for (size_t j = 0; j < m && !flip; ++j)
                if (skip[j]) flip = X.orientation(rt[j], v, Xk) != 1;
            if (flip) continue;
            const double ne = o.new_energy[k];
            if (ne >= o.energy || std::isinf(ne) || std::isnan(ne)) continue;
            taken = true;
            tiny = std::fabs(ne - o.energy) < 1e-5;
            pos = Xk;
        }
        if (taken && tiny) taken = false;
        if (!taken) {
            is_moved = step != 0;
            break;
        }
        is_moved = true;
    }
    p = pos;
    return is_moved;
}

// smoothSingle's (kind 0) or smoothSurface's (kind 1, 2) body after the
// try-to-round step, host-driven, for a ring with an unrounded vertex other
// than v; the same statuses and writes as a twg_smooth_batch problem
int VertexSmoother::smoothExact(int v, int kind, const std::vector<int>& ring)
{
    std::vector<std::array<int, 4>> rt;
    for (int t : ring) rt.push_back(tets[t]);
    // the validity test orients on posf whatever the rounding (:163-172)
    if (posfFlip(rt, -1, nullptr)) return TWG_SMOOTH_INVALID;
    Point_3f pn;
    if (!newtonExact(v, ring, pn)) return TWG_SMOOTH_NOT_MOVED;
    TetVertex& tv = tet_vertices[v];
    if (kind == 0) {
        if (isFlipAt(rt, v, &pn)) return TWG_SMOOTH_FLIP_END;
        tv.posf = pn;
        return TWG_SMOOTH_MOVED;
    }
    // projection (:354-362)
    Point_3f pf;
    (kind == 2 ? geo_b_tree : geo_sf_tree).squared_distance(pn, &pf);
    if (isFlipAt(rt, v, &pf)) return TWG_SMOOTH_REJECT_FLIP;
    double om = 0.0, nm = 0.0;
    for (int t : ring) om = std::max(om, tet_qualities[t].slim_energy);
    std::vector<double> nq;
    qualitiesAt(rt, v, pf, nq);
    for (double q : nq) nm = std::max(nm, q);
    if (!(nm < om)) return TWG_SMOOTH_REJECT_QUALITY;
    // the one-ring surface faces (:316-327) with v at pf (:403-429)
    std::vector<std::array<int, 3>> tri_ids;
    for (int t : ring)
        for (int j = 0; j < 4; ++j)
            if (tets[t][j] != v && is_surface_fs[t][j] != State::NOT_SURFACE) {
                std::array<int, 3> tri = {{tets[t][(j + 1) % 4], tets[t][(j + 2) % 4],
                                           tets[t][(j + 3) % 4]}};
                std::sort(tri.begin(), tri.end());
                tri_ids.push_back(tri);
            }
    std::sort(tri_ids.begin(), tri_ids.end());
    tri_ids.erase(std::unique(tri_ids.begin(), tri_ids.end()), tri_ids.end());
    std::vector<Triangle_3f> tris;
    for (const auto& tri : tri_ids) {
        int k = (int)(std::find(tri.begin(), tri.end(), v) - tri.begin());
        tris.push_back({{pf, tet_vertices[tri[(k + 1) % 3]].posf, tet_vertices[tri[(k + 2) % 3]].posf}});
    }
    std::vector<uint8_t> out;
    areFacesOutEnvelop(tris, out);
    for (uint8_t o : out)
        if (o) return TWG_SMOOTH_REJECT_ENVELOPE;
    tv.posf = pf;
    for (size_t i = 0; i < ring.size(); ++i) {
        TetQuality q;
        q.slim_energy = nq[i];
        tet_qualities[ring[i]] = q;
    }
    return TWG_SMOOTH_MOVED;
}

// the calTetQualities pass that ends smoothSingle
void VertexSmoother::finishSingle()
{
    std::vector<std::array<int, 4>> live;
    std::vector<int> ids;
    for (size_t i = 0; i < tets.size(); ++i)
        if (!t_is_removed[i]) {
            live.push_back(tets[i]);
            ids.push_back((int)i);
        }
    std::vector<TetQuality> qs;
    calTetQualities(live, qs);
    for (size_t k = 0; k < ids.size(); ++k) tet_qualities[ids[k]] = qs[k];
}

std::vector<int> VertexSmoother::surfaceCandidates()
{
    std::vector<int> verts;
    for (int v = 0; v < (int)tet_vertices.size(); ++v) {
        if (v_is_removed[v]) continue;
        TetVertex& tv = tet_vertices[v];
        if (!tv.is_on_surface) continue;
        if (tv.is_locked) continue;
        if (isIsolated(v)) {
            tv.is_on_surface = false;
            tv.is_on_boundary = false;
            continue;
        }
        if (!isBoundaryPoint(v)) tv.is_on_boundary = false;
        verts.push_back(v);
    }
    return verts;
}

std::vector<std::vector<int>> VertexSmoother::orderLevels(const std::vector<int>& verts) const
{
    std::vector<int> level(tet_vertices.size(), -1);
    std::vector<std::vector<int>> levels;
    for (int v : verts) {
        int l = 0;
        for (int t : tet_vertices[v].conn_tets)
            for (int j = 0; j < 4; ++j) {
                const int u = tets[t][j];
                if (u != v && level[u] >= l) l = level[u] + 1;
            }
        level[v] = l;
        if (l == (int)levels.size()) levels.emplace_back();
        levels[l].push_back(v);
    }
    return levels;
}

bool VertexSmoother::allRounded() const
{
    for (size_t v = 0; v < tet_vertices.size(); ++v)
        if (!v_is_removed[v] && !tet_vertices[v].is_rounded) return false;
    return true;
}

void VertexSmoother::runLevelsOnDevice(const std::vector<std::vector<int>>& levels, bool surface,
                                       int& suc)
{
    const size_t nv = tet_vertices.size(), nt = tets.size();
    std::vector<int32_t> center, ring_off{0}, ring_tets, ring_tid, tri_off{0}, tri_ab;
    std::vector<uint8_t> kind;
    std::vector<size_t> lv_start{0};
    std::vector<int> ring;
    for (const auto& lv : levels) {
        for (int v : lv) {
            const int k = surface ? (tet_vertices[v].is_on_boundary ? 2 : 1) : 0;
            center.push_back(v);
            kind.push_back((uint8_t)k);
            oneRing(v, ring);
            for (int t : ring) {
                for (int j = 0; j < 4; ++j) ring_tets.push_back(tets[t][j]);
                ring_tid.push_back(t);
            }
            ring_off.push_back((int32_t)(ring_tets.size() / 4));
            if (k != 0) {
                // one-ring surface faces, as runBatch lists them
                std::vector<std::array<int, 3>> tri_ids;
                for (int t : ring)
                    for (int j = 0; j < 4; ++j)
                        if (tets[t][j] != v && is_surface_fs[t][j] != State::NOT_SURFACE) {
                            std::array<int, 3> tri = {{tets[t][(j + 1) % 4], tets[t][(j + 2) % 4],
                                                       tets[t][(j + 3) % 4]}};
                            std::sort(tri.begin(), tri.end());
                            tri_ids.push_back(tri);
                        }
                std::sort(tri_ids.begin(), tri_ids.end());
                tri_ids.erase(std::unique(tri_ids.begin(), tri_ids.end()), tri_ids.end());
                for (const auto& tri : tri_ids) {
                    const int kk = (int)(std::find(tri.begin(), tri.end(), v) - tri.begin());
                    tri_ab.push_back(tri[(kk + 1) % 3]);
                    tri_ab.push_back(tri[(kk + 2) % 3]);
                }
            }
            tri_off.push_back((int32_t)(tri_ab.size() / 2));
        }
        lv_start.push_back(center.size());
    }
    const size_t np = center.size(), nr = ring_tid.size(), ntri = tri_ab.size() / 2;
    if (np == 0) return;
    std::vector<double> xyzq(3 * nv + nt);
    for (size_t v = 0; v < nv; ++v)
        for (int c = 0; c < 3; ++c) xyzq[c * nv + v] = tet_vertices[v].posf[c];
    for (size_t t = 0; t < nt; ++t)
        xyzq[3 * nv + t] = tet_qualities.size() > t ? tet_qualities[t].slim_energy : 0.0;
    // device copies (runBatch's buffer slots 0..5)
    double* dx = (double*)dev.buffer(0, 8 * (3 * nv + nt));
    double *dy = dx + nv, *dz = dy + nv, *dq = dz + nv;
    int32_t* dring = (int32_t*)dev.buffer(1, 16 * nr + 16);
    int32_t* dcenter = (int32_t*)dev.buffer(2, 4 * (np + 2 * (np + 1) + 2 * ntri + nr) + 64);
    int32_t* droff = dcenter + np;
    int32_t* dtoff = droff + (np + 1);
    int32_t* dtri = dtoff + (np + 1);
    int32_t* dtid = dtri + 2 * ntri;
    uint8_t* dkind = (uint8_t*)dev.buffer(3, 2 * np);
    uint8_t* dstatus = dkind + np;
    double* dold = (double*)dev.buffer(4, 16 * nr + 16);
    double* dnew = dold + nr;
    double* dpout = (double*)dev.buffer(5, 24 * np);
    dev.h2d(dx, xyzq.data(), 8 * xyzq.size());
    dev.h2d(dring, ring_tets.data(), 16 * nr);
    dev.h2d(dcenter, center.data(), 4 * np);
    dev.h2d(droff, ring_off.data(), 4 * (np + 1));
    dev.h2d(dtoff, tri_off.data(), 4 * (np + 1));
    dev.h2d(dtri, tri_ab.data(), 8 * ntri);
    dev.h2d(dtid, ring_tid.data(), 4 * nr);
    dev.h2d(dkind, kind.data(), np);
    for (size_t l = 0; l + 1 < lv_start.size(); ++l) {
        const size_t s0 = lv_start[l], n = lv_start[l + 1] - s0;
        const size_t r0 = ring_off[s0], r1 = ring_off[s0 + n];
        twg_smooth_desc d;
        std::memset(&d, 0, sizeof d);
        d.x = dx;
        d.y = dy;
        d.z = dz;
        d.nv = (int64_t)nv;
        d.n = (int64_t)n;
        d.center = dcenter + s0;
        d.ring_off = droff + s0;  // offsets into the whole ring arrays
        d.ring_tets = dring;
        d.n_ring = (int64_t)nr;
        d.kind = dkind + s0;
        d.old_q = dold;
        d.tri_off = dtoff + s0;
        d.tri_ab = dtri;
        d.n_tri = (int64_t)ntri;
        d.sf_tree = geo_sf_tree.bvh();
        d.b_tree = geo_b_tree.bvh();
        d.eps2 = state.eps_2;
        d.sampling_dist = state.sampling_dist > 0 ? state.sampling_dist : 1.0;
        d.p_out = dpout + 3 * s0;
        d.status = dstatus + s0;
        d.new_q = dnew;
        if (surface) {
            // this level's old qualities, as the earlier levels left them
            twg_smooth_desc g = d;
            g.old_q = dold + r0;
            g.n_ring = (int64_t)(r1 - r0);
            check(twg_smooth_gather_q(dev.ctx(), &g, dtid + r0, dq, (int64_t)nt), "twg_smooth_gather_q");
        }
        check(twg_smooth_batch(dev.ctx(), &d, precision, TWG_SMOOTH_TILE_SPLIT, 1), "twg_smooth_batch");
        check(twg_smooth_apply(dev.ctx(), &d, dtid, dx, dy, dz, surface ? dq : nullptr, (int64_t)nt),
              "twg_smooth_apply");
    }
    std::vector<uint8_t> status(np);
    dev.d2h(status.data(), dstatus, np);
    dev.d2h(xyzq.data(), dx, 8 * xyzq.size());
    for (size_t p = 0; p < np; ++p) {
        const int v = center[p];
        if (status[p] == TWG_SMOOTH_BAD_INPUT)
            throw std::runtime_error("twg_smooth_batch: bad input for vertex " + std::to_string(v));
        if (status[p] == TWG_SMOOTH_MOVED) {
            tet_vertices[v].posf = {{xyzq[v], xyzq[nv + v], xyzq[2 * nv + v]}};
            if (kind[p] != 0)
                for (int32_t r = ring_off[p]; r < ring_off[p + 1]; ++r) {
                    TetQuality q;
                    q.slim_energy = xyzq[3 * nv + ring_tid[r]];
                    tet_qualities[ring_tid[r]] = q;
                }
        }
        afterMove(v, status[p]);
        if (status[p] == TWG_SMOOTH_MOVED || (!surface && status[p] == TWG_SMOOTH_FLIP_END)) suc++;
    }
}

// src/tetwild/VertexSmoother.cpp:114-242 with the vertex order's results.
// Why the levels give them: the body for v reads only its one-ring (the
// positions, rounding and qualities of vertices sharing a tet with v) and
// writes only v and its ring tets.  When v runs, the vertex-order loop has
// updated exactly the earlier ring vertices; in level order those have lower
// levels (done) and the later ones higher levels (not started), and no two
// vertices of a level share a tet.
void VertexSmoother::smoothSingle()
{
    counter = 0;
    suc_counter = 0;
    std::vector<int> verts;
    for (int v = 0; v < (int)tet_vertices.size(); ++v) {
        if (v_is_removed[v]) continue;
        const TetVertex& tv = tet_vertices[v];
        if (tv.is_on_bbox) continue;
        if (state.eps != State::EPSILON_INFINITE && tv.is_on_surface) continue;
        if (tv.is_locked) continue;
        verts.push_back(v);
    }
    const auto levels = orderLevels(verts);
    last_levels = (int)levels.size();
    if (allRounded()) {
        runLevelsOnDevice(levels, false, suc_counter);
        counter += (int)verts.size();
    } else {
        for (const auto& lv : levels) {
            runClass(lv, false, suc_counter);
            counter += (int)lv.size();
        }
    }
    finishSingle();
}

// src/tetwild/VertexSmoother.cpp:244-463 with the vertex order's results:
// the flag updates touch no position and run first, in vertex order, as the
// loop interleaves them; then the levels, as in smoothSingle
void VertexSmoother::smoothSurface()
{
    const std::vector<int> verts = surfaceCandidates();
    const auto levels = orderLevels(verts);
    last_levels = (int)levels.size();
    if (allRounded()) {
        runLevelsOnDevice(levels, true, suc_counter);
        counter += (int)verts.size();
    } else {
        for (const auto& lv : levels) {
            runClass(lv, true, suc_counter);
            counter += (int)lv.size();
        }
    }
}

// src/tetwild/VertexSmoother.cpp:114-242 in vertex order, one call per vertex
void VertexSmoother::smoothSingleSerial()
{
    counter = 0;
    suc_counter = 0;
    for (int v = 0; v < (int)tet_vertices.size(); ++v) {
        if (v_is_removed[v]) continue;
        const TetVertex& tv = tet_vertices[v];
        if (tv.is_on_bbox) continue;
        if (state.eps != State::EPSILON_INFINITE && tv.is_on_surface) continue;
        if (tv.is_locked) continue;
        counter++;
        int s = smoothOne(v, 0);
        if (s == TWG_SMOOTH_MOVED || s == TWG_SMOOTH_FLIP_END) suc_counter++;
    }
    finishSingle();
}

// src/tetwild/VertexSmoother.cpp:244-463 in vertex order, one call per vertex
void VertexSmoother::smoothSurfaceSerial()
{
    for (int v = 0; v < (int)tet_vertices.size(); ++v) {
        if (v_is_removed[v]) continue;
        TetVertex& tv = tet_vertices[v];
        if (!tv.is_on_surface) continue;
        if (tv.is_locked) continue;
        if (isIsolated(v)) {
            tv.is_on_surface = false;
            tv.is_on_boundary = false;
            continue;
        }
        if (!isBoundaryPoint(v)) tv.is_on_boundary = false;
        counter++;
        int s = smoothOne(v, tv.is_on_boundary ? 2 : 1);
        if (s == TWG_SMOOTH_MOVED) suc_counter++;
    }
}

// src/tetwild/VertexSmoother.cpp:19-46
void VertexSmoother::smooth()
{
    smoothSingle();
    if (state.eps >= 0) smoothSurface();
}

std::vector<std::vector<int>> VertexSmoother::colourClasses(const std::vector<int>& verts) const
{
    std::vector<int> colour(tet_vertices.size(), -1);
    std::vector<std::vector<int>> classes;
    std::vector<char> used;
    for (int v : verts) {
        used.assign(classes.size() + 1, 0);
        for (int t : tet_vertices[v].conn_tets)
            for (int j = 0; j < 4; ++j) {
                int u = tets[t][j];
                if (u != v && colour[u] >= 0) used[colour[u]] = 1;
            }
        int c = 0;
        while (used[c]) ++c;
        colour[v] = c;
        if (c == (int)classes.size()) classes.emplace_back();
        classes[c].push_back(v);
    }
    return classes;
}

// One colour class: the try-to-round step per vertex, then one batch for the
// rings the device decides alone and the host-driven path for the others.
// The vertices of a class share no tet, so the order inside it is free.
void VertexSmoother::runClass(const std::vector<int>& cls, bool surface, int& suc)
{
    std::vector<Problem> probs;
    std::vector<int> exact_v;
    for (int v : cls) {
        std::vector<int> ring;
        oneRing(v, ring);
        std::vector<std::array<int, 4>> rt;
        for (int t : ring) rt.push_back(tets[t]);
        tryRound(v, rt);
        bool needs_exact = false;
        for (const auto& t : rt) needs_exact = needs_exact || hasUnroundedOther(t, v);
        const int kind = surface ? (tet_vertices[v].is_on_boundary ? 2 : 1) : 0;
        if (needs_exact) exact_v.push_back(v);
        else probs.push_back(Problem{v, kind});
    }
    std::vector<uint8_t> st;
    runBatch(probs, st);
    auto count = [&](int s) {
        if (s == TWG_SMOOTH_MOVED || (!surface && s == TWG_SMOOTH_FLIP_END)) suc++;
    };
    for (size_t i = 0; i < probs.size(); ++i) {
        afterMove(probs[i].v, st[i]);
        count(st[i]);
    }
    for (int v : exact_v) {
        std::vector<int> ring;
        oneRing(v, ring);
        int s = smoothExact(v, surface ? (tet_vertices[v].is_on_boundary ? 2 : 1) : 0, ring);
        afterMove(v, s);
        count(s);
    }
}

void VertexSmoother::smoothSingleBatched()
{
    counter = 0;
    suc_counter = 0;
    std::vector<int> verts;
    for (int v = 0; v < (int)tet_vertices.size(); ++v) {
        if (v_is_removed[v]) continue;
        const TetVertex& tv = tet_vertices[v];
        if (tv.is_on_bbox || tv.is_locked) continue;
        if (state.eps != State::EPSILON_INFINITE && tv.is_on_surface) continue;
        verts.push_back(v);
    }
    for (const auto& cls : colourClasses(verts)) {
        runClass(cls, false, suc_counter);
        counter += (int)cls.size();
    }
    finishSingle();
}

void VertexSmoother::smoothSurfaceBatched()
{
    const std::vector<int> verts = surfaceCandidates();
    for (const auto& cls : colourClasses(verts)) {
        runClass(cls, true, suc_counter);
        counter += (int)cls.size();
    }
}

// ---- EdgeCollapser ---------------------------------------------------------
void EdgeCollapser::oneRing(int v_id, std::vector<int>& t_ids) const
{
    t_ids.assign(tet_vertices[v_id].conn_tets.begin(), tet_vertices[v_id].conn_tets.end());
    std::sort(t_ids.begin(), t_ids.end());
}

int EdgeCollapser::checkCollapse(int v1_id, int v2_id, std::vector<TetQuality>* new_qs)
{
    std::vector<int> codes;
    std::vector<std::vector<TetQuality>> qs;
    checkCollapses({{{v1_id, v2_id}}}, codes, new_qs ? &qs : nullptr);
    if (new_qs) *new_qs = qs[0];
    return codes[0];
}

// collapseAnEdge's checks (src/tetwild/EdgeCollapser.cpp:284-446) after its
// flag updates, in the reference's order
int EdgeCollapser::checkCollapseExact(int v1_id, int v2_id, std::vector<TetQuality>* new_qs)
{
    const TetVertex &a = tet_vertices[v1_id], &b = tet_vertices[v2_id];
    const bool degenerate = a.posf == b.posf;  // the edge length is 0
    if (a.is_on_boundary && !b.is_on_boundary && !degenerate && isPointOutBoundaryEnvelop(b.posf))
        return ENVELOP;
    if (a.is_on_surface && !b.is_on_surface && !degenerate && isPointOutEnvelop(b.posf))
        return ENVELOP;
    std::vector<int> ring;
    oneRing(v1_id, ring);
    std::vector<std::array<int, 4>> new_tets;
    for (int t : ring) {
        if (contains(tets[t], v2_id)) continue;
        std::array<int, 4> nt = tets[t];
        replace_id(nt, v1_id, v2_id);
        new_tets.push_back(nt);
    }
    if (isFlip(new_tets)) return FLIP;
    std::vector<TetQuality> qs;
    calTetQualities(new_tets, qs);
    if (energy_type != State::ENERGY_NA && is_check_quality) {
        TetQuality old_tq, new_tq;
        getCheckQuality(ring, old_tq);
        getCheckQuality(qs, new_tq);
        if (is_soft) old_tq.slim_energy = soft_energy;
        if (!a.is_rounded) new_tq.slim_energy = 0;  // an unroundable vertex goes anyway (:408)
        if (!degenerate && !new_tq.isBetterOrEqualThan(old_tq)) return QUALITY;
    }
    int code = SUCCESS;
    if (state.eps != State::EPSILON_NA && state.eps != State::EPSILON_INFINITE && a.is_on_surface) {
        // isCollapsable_epsilon (:727-770)
        std::vector<std::array<int, 3>> tri_ids;
        for (int t : ring)
            for (int j = 0; j < 4; ++j)
                if (tets[t][j] != v1_id && is_surface_fs[t][j] != State::NOT_SURFACE) {
                    std::array<int, 3> f = {{tets[t][(j + 1) % 4], tets[t][(j + 2) % 4],
                                             tets[t][(j + 3) % 4]}};
                    std::sort(f.begin(), f.end());
                    tri_ids.push_back(f);
                }
        std::sort(tri_ids.begin(), tri_ids.end());
        tri_ids.erase(std::unique(tri_ids.begin(), tri_ids.end()), tri_ids.end());
        std::vector<Triangle_3f> tris;
        for (auto f : tri_ids) {
            if (std::find(f.begin(), f.end(), v2_id) != f.end()) continue;
            *std::find(f.begin(), f.end(), v1_id) = v2_id;
            tris.push_back({{tet_vertices[f[0]].posf, tet_vertices[f[1]].posf, tet_vertices[f[2]].posf}});
        }
        std::vector<uint8_t> out;
        areFacesOutEnvelop(tris, out);
        if (!degenerate)
            for (uint8_t o : out)
                if (o) return ENVELOP;
        code = ENVELOP_SUC;
    }
    if (new_qs) *new_qs = qs;
    return code;
}

void EdgeCollapser::checkCollapses(const std::vector<std::array<int, 2>>& edges,
                                   std::vector<int>& codes,
                                   std::vector<std::vector<TetQuality>>* new_qs)
{
    const size_t n = edges.size();
    codes.assign(n, SUCCESS);
    if (new_qs) new_qs->assign(n, {});
    if (n == 0) return;
    // the flag updates at the top of collapseAnEdge, in candidate order
    for (const auto& e : edges) {
        TetVertex& v1 = tet_vertices[e[0]];
        if (v1.is_on_surface && isIsolated(e[0])) {
            v1.is_on_surface = false;
            v1.is_on_boundary = false;
        }
        if (!isBoundaryPoint(e[0])) v1.is_on_boundary = false;
    }
    if (stamp_.size() < tet_vertices.size()) stamp_.assign(tet_vertices.size(), -1);
    std::vector<int> global;
    auto local = [&](int v) {
        if (stamp_[v] < 0) {
            stamp_[v] = (int)global.size();
            global.push_back(v);
        }
        return stamp_[v];
    };
    std::vector<int32_t> v1s(n), v2s(n), ring_off(n + 1), tri_off(n + 1), ring_tets, tri_ids;
    std::vector<uint8_t> flags(n);
    std::vector<double> old_q;
    std::vector<std::vector<int>> rings(n);
    ring_off[0] = tri_off[0] = 0;
    for (size_t c = 0; c < n; ++c) {
        const int v1 = edges[c][0], v2 = edges[c][1];
        v1s[c] = local(v1);
        v2s[c] = local(v2);
        const TetVertex &a = tet_vertices[v1], &b = tet_vertices[v2];
        flags[c] = (uint8_t)((a.is_on_surface ? TWG_CV1_SURFACE : 0) |
                             (a.is_on_boundary ? TWG_CV1_BOUNDARY : 0) |
                             (b.is_on_surface ? TWG_CV2_SURFACE : 0) |
                             (b.is_on_boundary ? TWG_CV2_BOUNDARY : 0) |
                             (a.is_rounded ? 0 : TWG_CV1_UNROUNDED));
        oneRing(v1, rings[c]);
        for (int t : rings[c]) {
            for (int j = 0; j < 4; ++j) ring_tets.push_back(local(tets[t][j]));
            old_q.push_back(tet_qualities[t].slim_energy);
        }
        ring_off[c + 1] = (int32_t)(ring_tets.size() / 4);
        // isCollapsable_epsilon's triangles, src/tetwild/EdgeCollapser.cpp:742-764
        std::vector<std::array<int, 3>> tri;
        for (int t : rings[c])
            for (int j = 0; j < 4; ++j)
                if (tets[t][j] != v1 && is_surface_fs[t][j] != State::NOT_SURFACE) {
                    std::array<int, 3> f = {{tets[t][(j + 1) % 4], tets[t][(j + 2) % 4],
                                             tets[t][(j + 3) % 4]}};
                    std::sort(f.begin(), f.end());
                    tri.push_back(f);
                }
        std::sort(tri.begin(), tri.end());
        tri.erase(std::unique(tri.begin(), tri.end()), tri.end());
        for (auto f : tri) {
            if (std::find(f.begin(), f.end(), v2) != f.end()) continue;
            *std::find(f.begin(), f.end(), v1) = v2;
            for (int k = 0; k < 3; ++k) tri_ids.push_back(local(f[k]));
        }
        tri_off[c + 1] = (int32_t)(tri_ids.size() / 3);
    }
    const size_t nv = global.size(), nr = ring_tets.size() / 4, nt = tri_ids.size() / 3;
    std::vector<double> x(nv), y(nv), z(nv);
    for (size_t k = 0; k < nv; ++k) {
        const Point_3f& p = tet_vertices[global[k]].posf;
        x[k] = p[0];
        y[k] = p[1];
        z[k] = p[2];
        stamp_[global[k]] = -1;
    }
    char* b0 = (char*)dev.buffer(0, 24 * nv);
    double *dx = (double*)b0, *dy = dx + nv, *dz = dy + nv;
    int32_t* dring = (int32_t*)dev.buffer(1, 16 * nr + 16);
    char* b2 = (char*)dev.buffer(2, 4 * (4 * n + 2) + 12 * nt + 64);
    int32_t* dv1 = (int32_t*)b2;
    int32_t* dv2 = dv1 + n;
    int32_t* droff = dv2 + n;
    int32_t* dtoff = droff + (n + 1);
    int32_t* dtri = dtoff + (n + 1);
    uint8_t* dflags = (uint8_t*)dev.buffer(3, n + n);
    uint8_t* dstatus = dflags + n;
    double* dq = (double*)dev.buffer(4, 16 * nr + 16);
    double* dnewq = dq + nr;
    dev.h2d(dx, x.data(), 8 * nv);
    dev.h2d(dy, y.data(), 8 * nv);
    dev.h2d(dz, z.data(), 8 * nv);
    dev.h2d(dring, ring_tets.data(), 16 * nr);
    dev.h2d(dv1, v1s.data(), 4 * n);
    dev.h2d(dv2, v2s.data(), 4 * n);
    dev.h2d(droff, ring_off.data(), 4 * (n + 1));
    dev.h2d(dtoff, tri_off.data(), 4 * (n + 1));
    dev.h2d(dtri, tri_ids.data(), 12 * nt);
    dev.h2d(dflags, flags.data(), n);
    dev.h2d(dq, old_q.data(), 8 * nr);
    twg_collapse_desc d;
    std::memset(&d, 0, sizeof d);
    d.x = dx;
    d.y = dy;
    d.z = dz;
    d.nv = (int64_t)nv;
    d.n = (int64_t)n;
    d.v1 = dv1;
    d.v2 = dv2;
    d.flags = dflags;
    d.ring_off = droff;
    d.ring_tets = dring;
    d.n_ring = (int64_t)nr;
    d.old_q = dq;
    d.tri_off = dtoff;
    d.tri_ids = dtri;
    d.n_tri = (int64_t)nt;
    d.sf_tree = geo_sf_tree.bvh();
    d.b_tree = geo_b_tree.bvh();
    d.eps2 = state.eps_2;
    d.sampling_dist = state.sampling_dist > 0 ? state.sampling_dist : 1.0;
    d.check_quality = energy_type != State::ENERGY_NA && is_check_quality;
    d.check_eps = state.eps != State::EPSILON_NA && state.eps != State::EPSILON_INFINITE;
    d.is_soft = is_soft;
    d.soft_energy = soft_energy;
    d.status = dstatus;
    d.new_q = dnewq;
    check(twg_collapse_batch(dev.ctx(), &d, precision, 0, 0), "twg_collapse_batch");
    std::vector<uint8_t> st(n);
    std::vector<double> newq(nr);
    dev.d2h(st.data(), dstatus, n);
    dev.d2h(newq.data(), dnewq, 8 * nr);
    for (size_t c = 0; c < n; ++c) {
        if (st[c] == TWG_COLLAPSE_BAD_INPUT)
            throw std::runtime_error("twg_collapse_batch: bad input for edge " +
                                     std::to_string(edges[c][0]) + "-" + std::to_string(edges[c][1]));
        // a new tet with an unrounded vertex: the flip test needs ExactPositions
        bool needs_exact = false;
        for (int t : rings[c]) {
            if (contains(tets[t], edges[c][1])) continue;
            std::array<int, 4> nt = tets[t];
            replace_id(nt, edges[c][0], edges[c][1]);
            needs_exact = needs_exact || hasUnroundedOther(nt, -1);
        }
        if (needs_exact) {
            codes[c] = checkCollapseExact(edges[c][0], edges[c][1], new_qs ? &(*new_qs)[c] : nullptr);
            continue;
        }
        codes[c] = st[c];
        if (!new_qs || (st[c] != SUCCESS && st[c] != ENVELOP_SUC)) continue;
        for (size_t i = 0; i < rings[c].size(); ++i) {
            const auto& t = tets[rings[c][i]];
            if (std::find(t.begin(), t.end(), edges[c][1]) != t.end()) continue;  // removed
            TetQuality q;
            q.slim_energy = newq[ring_off[c] + i];
            (*new_qs)[c].push_back(q);
        }
    }
}

// ---- EdgeRemover -----------------------------------------------------------
namespace {

// the two vertices of every ring tet other than v1, v2, chained around the
// edge (n12_v_ids) and the ring positions in chain order (n12_t_ids),
// src/tetwild/EdgeRemover.cpp:276-318
void chain_ring(const std::vector<std::array<int, 4>>& ring, int v1, int v2, std::vector<int>& nv,
                std::vector<int>& nt)
{
    const int m = (int)ring.size();
    std::vector<std::array<int, 2>> es(m);
    for (int i = 0; i < m; ++i) {
        int c = 0;
        for (int v : ring[i])
            if (v != v1 && v != v2 && c < 2) es[i][c++] = v;
    }
    std::vector<bool> seen(m, false);
    seen[0] = true;
    nv = {es[0][0], es[0][1]};
    nt = {0};
    for (int i = 0; i < m - 2; ++i)
        for (int j = 0; j < m; ++j) {
            if (seen[j]) continue;
            if (es[j][0] == nv.back()) {
                seen[j] = true;
                nv.push_back(es[j][1]);
            } else if (es[j][1] == nv.back()) {
                seen[j] = true;
                nv.push_back(es[j][0]);
            }
            if (seen[j]) {
                nt.push_back(j);
                break;
            }
        }
    for (int j = 0; j < m; ++j)
        if (!seen[j]) {
            nt.push_back(j);
            break;
        }
}

}  // namespace

EdgeRemover::Removal EdgeRemover::checkRemoval(int v1_id, int v2_id, const std::vector<int>& old_t_ids)
{
    std::vector<Removal> out;
    checkRemovals({{{v1_id, v2_id}}}, {old_t_ids}, out);
    return out[0];
}

void EdgeRemover::checkRemovals(const std::vector<std::array<int, 2>>& edges,
                                const std::vector<std::vector<int>>& rings,
                                std::vector<Removal>& out)
{
    const size_t n = edges.size();
    out.assign(n, Removal());
    // configurations of every edge, one group each
    std::vector<std::array<int, 4>> gt;   // group tets (global ids)
    std::vector<int32_t> goff{0};
    std::vector<int> first_group(n, -1);
    std::vector<std::vector<int>> chain_v(n), chain_t(n);
    auto close_group = [&]() { goff.push_back((int32_t)gt.size()); };
    for (size_t e = 0; e < n; ++e) {
        const int v1 = edges[e][0], v2 = edges[e][1];
        const auto& ids = rings[e];
        const int m = (int)ids.size();
        if (m < 3 || m > 5) continue;
        std::vector<std::array<int, 4>> ring(m);
        for (int i = 0; i < m; ++i) ring[i] = tets[ids[i]];
        first_group[e] = (int)goff.size() - 1;
        if (m == 3) {  // src/tetwild/EdgeRemover.cpp:127-157
            std::array<int, 2> vi{};
            int c = 0;
            for (int v : ring[0])
                if (v != v1 && v != v2) vi[c++] = v;
            int a = contains(ring[1], vi[0]) ? 1 : 2, b = 3 - a;
            std::array<int, 4> t0 = ring[a], t1 = ring[b];
            replace_id(t0, v1, vi[1]);
            replace_id(t1, v2, vi[0]);
            gt.push_back(t0);
            gt.push_back(t1);
            close_group();
        } else if (m == 4) {  // :320-345, two rotations
            chain_ring(ring, v1, v2, chain_v[e], chain_t[e]);
            for (int i = 0; i < 2; ++i) {
                const int a = chain_v[e][i], b = chain_v[e][2 + i];
                for (int j = 0; j < 4; ++j) {
                    std::array<int, 4> t = ring[j];
                    if (contains(t, a)) replace_id(t, v2, b);
                    else replace_id(t, v1, a);
                    gt.push_back(t);
                }
                close_group();
            }
        } else {  // :458-575: pair k (around n12 vertex k), then middle pair k
            chain_ring(ring, v1, v2, chain_v[e], chain_t[e]);
            const auto& nv = chain_v[e];
            const auto& nt = chain_t[e];
            for (int k = 0; k < 5; ++k) {
                std::array<int, 4> t0 = ring[nt[k]], t1 = ring[nt[k]];
                replace_id(t0, v1, nv[(k + 4) % 5]);
                replace_id(t1, v2, nv[(k + 4) % 5]);
                gt.push_back(t0);
                gt.push_back(t1);
                close_group();
            }
            for (int k = 0; k < 5; ++k) {
                std::array<int, 4> t0 = ring[nt[(k + 2) % 5]], t1 = ring[nt[(k + 2) % 5]];
                replace_id(t0, v1, nv[k]);
                replace_id(t1, v2, nv[k]);
                gt.push_back(t0);
                gt.push_back(t1);
                close_group();
            }
        }
    }
    const size_t ng = goff.size() - 1, ntt = gt.size();
    std::vector<uint8_t> gflip(ng);
    std::vector<double> gmax(ng), gq(ntt);
    if (ng > 0) {
        // compact vertex set, upload, one launch
        if (stamp_.size() < tet_vertices.size()) stamp_.assign(tet_vertices.size(), -1);
        std::vector<int> global;
        std::vector<int32_t> lt(4 * ntt);
        for (size_t k = 0; k < ntt; ++k)
            for (int j = 0; j < 4; ++j) {
                int v = gt[k][j];
                if (stamp_[v] < 0) {
                    stamp_[v] = (int)global.size();
                    global.push_back(v);
                }
                lt[4 * k + j] = stamp_[v];
            }
        const size_t nv = global.size();
        std::vector<double> x(nv), y(nv), z(nv);
        for (size_t k = 0; k < nv; ++k) {
            const Point_3f& p = tet_vertices[global[k]].posf;
            x[k] = p[0];
            y[k] = p[1];
            z[k] = p[2];
            stamp_[global[k]] = -1;
        }
        double* dx = (double*)dev.buffer(0, 24 * nv);
        double *dy = dx + nv, *dz = dy + nv;
        int32_t* dt = (int32_t*)dev.buffer(1, 16 * ntt + 16);
        int32_t* doff = (int32_t*)dev.buffer(2, 4 * (ng + 1));
        uint8_t* dfl = (uint8_t*)dev.buffer(3, ng);
        double* dmax = (double*)dev.buffer(4, 8 * (ng + ntt));
        double* dq = dmax + ng;
        dev.h2d(dx, x.data(), 8 * nv);
        dev.h2d(dy, y.data(), 8 * nv);
        dev.h2d(dz, z.data(), 8 * nv);
        dev.h2d(dt, lt.data(), 16 * ntt);
        dev.h2d(doff, goff.data(), 4 * (ng + 1));
        check(twg_tet_groups(dev.ctx(), dt, doff, (int64_t)ng, (int64_t)ntt, dx, dy, dz,
                             (int64_t)nv, dfl, dmax, dq),
              "twg_tet_groups");
        dev.d2h(gflip.data(), dfl, ng);
        dev.d2h(gmax.data(), dmax, 8 * ng);
        dev.d2h(gq.data(), dq, 8 * ntt);
        for (size_t g = 0; g < ng; ++g)
            if (gflip[g] == TWG_GROUP_BAD_INPUT) throw std::runtime_error("twg_tet_groups: bad input");
        // a group with an unrounded vertex: its flip test through ExactPositions
        for (size_t g = 0; g < ng; ++g) {
            bool ex = false;
            for (int k = goff[g]; k < goff[g + 1] && !ex; ++k) ex = hasUnroundedOther(gt[k], -1);
            if (!ex) continue;
            std::vector<std::array<int, 4>> ts(gt.begin() + goff[g], gt.begin() + goff[g + 1]);
            gflip[g] = isFlipAt(ts, -1, nullptr) ? 1 : 0;
        }
    }
    // the reference's decisions, edge by edge
    auto better = [&](double nq, double oq) {
        if (equal_buget > 0) {
            equal_buget--;
            return nq <= oq;
        }
        return nq < oq;
    };
    auto take = [&](Removal& r, int g) {
        for (int k = goff[g]; k < goff[g + 1]; ++k) {
            r.new_tets.push_back(gt[k]);
            TetQuality q;
            q.slim_energy = gq[k];
            r.new_qs.push_back(q);
        }
    };
    for (size_t e = 0; e < n; ++e) {
        if (first_group[e] < 0) continue;
        const auto& ids = rings[e];
        const int m = (int)ids.size(), g0 = first_group[e];
        double old_max = 0;
        for (int t : ids)
            if (tet_qualities[t].slim_energy > old_max) old_max = tet_qualities[t].slim_energy;
        Removal& r = out[e];
        if (m == 3) {
            if (gflip[g0] || !better(gmax[g0], old_max)) continue;
            r.kind = 3;
            take(r, g0);
        } else if (m == 4) {
            int best = -1;
            bool fail = false;
            for (int i = 0; i < 2 && !fail; ++i) {
                if (gflip[g0 + i]) continue;
                if (!better(gmax[g0 + i], old_max)) {
                    fail = true;
                    break;
                }
                old_max = gmax[g0 + i];
                best = i;
            }
            if (fail || best < 0) continue;
            r.kind = 4;
            take(r, g0 + best);
        } else {
            bool ok[5] = {true, true, true, true, true};
            for (int k = 0; k < 5; ++k) {
                if (!ok[(k + 1) % 5] && !ok[(k + 4) % 5]) continue;
                if (gflip[g0 + k]) {
                    ok[(k + 1) % 5] = false;
                    ok[(k + 4) % 5] = false;
                }
            }
            int sel = -1;
            for (int i = 0; i < 5; ++i) {
                if (!ok[i] || gflip[g0 + 5 + i]) continue;
                const int l = (i + 1) % 5, rr = (i + 4) % 5;
                double mx = std::max(gmax[g0 + 5 + i], std::max(gmax[g0 + l], gmax[g0 + rr]));
                if (!better(mx, old_max)) continue;
                old_max = mx;
                sel = i;
            }
            if (sel < 0) continue;
            r.kind = 5;
            take(r, g0 + (sel + 1) % 5);
            take(r, g0 + (sel + 4) % 5);
            take(r, g0 + 5 + sel);
        }
    }
}

}  // namespace tetwild_amd
