// tetwild_shim.h -- C++ host classes with TetWild's LocalOperations /
// VertexSmoother call surface, backed by libtetwild_amd.so (the MI355X
// kernels, include/tetwild_gpu.h).
//
// The reference classes (src/tetwild/LocalOperations.h:33-125,
// src/tetwild/VertexSmoother.h:19-51) hold references to the mesh arrays
// owned by MeshRefinement and use CGAL/geogram types.  These mirrors keep the
// method names, argument meaning and error conventions (MAX_ENERGY sentinels,
// bool results) over std:: types that are layout-compatible with the
// reference's double-precision points:
//   Point_3f     -> std::array<double, 3>       (CGAL Epick Point_3)
//   Triangle_3f  -> std::array<Point_3f, 3>
//   GEO::MeshFacetsAABBWithEps -> tetwild_amd::Envelope (a device AABB tree)
// Every call gathers the vertices it touches into compact device arrays, so
// the GPU always sees the host mesh as it is at the time of the call (the
// driver may mutate the mesh freely between calls).
//
// Unrounded vertices (is_rounded false) have an exact rational position
// (TetVertex::pos, a CGAL Epeck Point_3, in the reference) that this mirror
// does not hold.  The reference orients a tet on pos whenever one of its
// vertices is unrounded (isTetFlip, src/tetwild/LocalOperations.cpp:535-553);
// the mirrors send exactly those tests to the caller's ExactPositions and
// decide every other one on the device.  A test that needs the exact
// positions with no ExactPositions set throws std::logic_error rather than
// falling back to posf.  For a rounded vertex pos == Point_3(posf), as the
// reference keeps it, so posf stands for pos there.
#pragma once

#include <array>
#include <climits>
#include <cstdint>
#include <limits>
#include <memory>
#include <unordered_set>
#include <vector>

#include "../../include/tetwild_gpu.h"

namespace tetwild_amd {

using Point_3f = std::array<double, 3>;
using Triangle_3f = std::array<Point_3f, 3>;

// TetVertex (src/tetwild/TetmeshElements.h:24-64), double-precision part
struct TetVertex {
    Point_3f posf{{0, 0, 0}};
    bool is_rounded = true;
    bool is_on_surface = false;
    bool is_on_bbox = false;
    bool is_on_boundary = false;
    bool is_locked = false;
    std::unordered_set<int> conn_tets;
};

// TetQuality (src/tetwild/TetmeshElements.h:66-111)
struct TetQuality {
    double min_d_angle = 0;
    double max_d_angle = 0;
    double slim_energy = 0;
    double volume = 0;
    bool isBetterThan(const TetQuality& tq) const { return slim_energy < tq.slim_energy; }
    bool isBetterOrEqualThan(const TetQuality& tq) const { return slim_energy <= tq.slim_energy; }
};

// The parts of State (src/tetwild/State.h:22-60) the path reads
struct State {
    static constexpr double MAX_ENERGY = 1e50;
    static constexpr int NOT_SURFACE = INT_MAX;
    static constexpr int EPSILON_INFINITE = -2;
    static constexpr int EPSILON_NA = -1;
    static constexpr int ENERGY_NA = 0;
    static constexpr int ENERGY_AMIPS = 2;
    double eps = 0;
    double eps_2 = 0;
    double sampling_dist = 0;
    bool is_mesh_closed = true;
    bool use_energy_max = true;
    // Args::filter_energy_thres (include/tetwild/Args.h:63), read by getFilterEnergy
    double filter_energy_thres = 10;
};

// The caller's exact positions of the unrounded vertices.
class ExactPositions {
public:
    virtual ~ExactPositions() = default;
    // CGAL::orientation(pos[t[0]], pos[t[1]], pos[t[2]], pos[t[3]]) as +1
    // (POSITIVE), 0 or -1, where pos of a rounded vertex is its posf and,
    // when moved >= 0, vertex `moved` sits at the double point `at`.
    virtual int orientation(const std::array<int, 4>& t, int moved, const Point_3f& at) = 0;
    // v has just become rounded: pos[v] = Point_3(posf[v]) from now on
    virtual void rounded(int v) { (void)v; }
};

// GPU context shared by the objects of one mesh (one device, one stream).
class Device {
public:
    explicit Device(int device = 0);
    ~Device();
    Device(const Device&) = delete;
    Device& operator=(const Device&) = delete;
    twg_ctx* ctx() const { return ctx_; }
    // growable device buffer i (0..7), at least `bytes` long
    void* buffer(int i, size_t bytes);
    void h2d(void* dst, const void* src, size_t bytes);
    void d2h(void* dst, const void* src, size_t bytes);
    void sync();

private:
    twg_ctx* ctx_ = nullptr;
    void* buf_[8] = {nullptr};
    size_t cap_[8] = {0};
};

// Process-wide device used by the static AMIPS functions.
Device& default_device();

// Stand-in for GEO::MeshFacetsAABBWithEps over a triangle surface: a device
// tree built once (MeshRefinement builds its trees once per refine,
// src/tetwild/MeshRefinement.cpp:209-213).
class Envelope {
public:
    Envelope(Device& dev, const std::vector<Point_3f>& V, const std::vector<std::array<int, 3>>& F);
    ~Envelope();
    Envelope(const Envelope&) = delete;
    Envelope& operator=(const Envelope&) = delete;
    const twg_bvh* bvh() const { return bvh_; }
    // squared distance to the surface and the nearest point
    double squared_distance(const Point_3f& p, Point_3f* nearest = nullptr) const;

private:
    Device& dev_;
    twg_bvh* bvh_ = nullptr;
};

class LocalOperations {
public:
    std::vector<TetVertex>& tet_vertices;
    std::vector<std::array<int, 4>>& tets;
    std::vector<std::array<int, 4>>& is_surface_fs;
    std::vector<bool>& v_is_removed;
    std::vector<bool>& t_is_removed;
    std::vector<TetQuality>& tet_qualities;
    const Envelope& geo_sf_tree;
    const Envelope& geo_b_tree;
    State& state;
    Device& dev;
    int counter = 0;
    int suc_counter = 0;
    // TWG_FP64 (default) or TWG_MIXED: results are identical, mixed screens in FP32
    int precision = TWG_FP64;
    // exact orientations for tets with unrounded vertices (may stay null on a
    // mesh whose vertices are all rounded)
    ExactPositions* exact = nullptr;

    LocalOperations(std::vector<TetVertex>& t_vs, std::vector<std::array<int, 4>>& ts,
                    std::vector<std::array<int, 4>>& is_sf_fs, std::vector<bool>& v_is_rm,
                    std::vector<bool>& t_is_rm, std::vector<TetQuality>& tet_qs,
                    const Envelope& sf_tree, const Envelope& b_tree, State& st, Device& device)
        : tet_vertices(t_vs), tets(ts), is_surface_fs(is_sf_fs), v_is_removed(v_is_rm),
          t_is_removed(t_is_rm), tet_qualities(tet_qs), geo_sf_tree(sf_tree), geo_b_tree(b_tree),
          state(st), dev(device)
    {
    }

    // src/tetwild/LocalOperations.h:69-111
    void calTetQualities(const std::vector<std::array<int, 4>>& new_tets,
                         std::vector<TetQuality>& tet_qs, bool all_measure = false);
    bool isFlip(const std::vector<std::array<int, 4>>& new_tets);
    bool isTetFlip(const std::array<int, 4>& t);
    bool isTetFlip(int t_id) { return isTetFlip(tets[t_id]); }
    void getCheckQuality(const std::vector<TetQuality>& tet_qs, TetQuality& tq);
    void getCheckQuality(const std::vector<int>& t_ids, TetQuality& tq);
    void getAvgMaxEnergy(double& avg_tq, double& max_tq);
    double getMaxEnergy();
    double getSecondMaxEnergy(double max_energy);
    double getFilterEnergy(bool& is_clean_up);
    bool isFaceOutEnvelop(const Triangle_3f& tri);
    bool isFaceOutEnvelop_sampling(const Triangle_3f& tri) { return isFaceOutEnvelop(tri); }
    bool isPointOutEnvelop(const Point_3f& p);
    bool isPointOutBoundaryEnvelop(const Point_3f& p);
    bool isIsolated(int v_id);
    bool isBoundaryPoint(int v_id);
    bool isEdgeOnBoundary(int v1_id, int v2_id);
    // batched forms of the envelope tests (one launch for many queries)
    void areFacesOutEnvelop(const std::vector<Triangle_3f>& tris, std::vector<uint8_t>& out);
    void arePointsOutEnvelop(const std::vector<Point_3f>& ps, std::vector<uint8_t>& out);

    static double comformalAMIPSEnergy_new(const double* T);
    static void comformalAMIPSJacobian_new(const double* T, double* result_0);
    static void comformalAMIPSHessian_new(const double* T, double* result_0);

protected:
    // sums, maxima and histogram of the stored slim energies (one device pass)
    twg_energy_summary energySummary();
    // compact gather: tets -> local ids, x/y/z of the touched vertices
    struct Gathered {
        std::vector<int32_t> tets;
        std::vector<double> x, y, z;
        std::vector<int> global;  // local -> global vertex id
    };
    // ov >= 0: vertex ov is gathered at *op instead of its posf
    void gather(const std::vector<std::array<int, 4>>& ts, Gathered& g, int pin = -1, int ov = -1,
                const Point_3f* op = nullptr);
    std::vector<int> stamp_;  // per-vertex local id (or -1), sized lazily

    // isFlip with vertex v at *at (v < 0: no vertex moved), the reference's
    // rule per tet: on posf when every vertex other than the moved one is
    // rounded (the moved one sits at a double point, so that is its exact
    // orientation), else through ExactPositions
    bool isFlipAt(const std::vector<std::array<int, 4>>& ts, int v, const Point_3f* at);
    // some vertex of t other than v is unrounded
    bool hasUnroundedOther(const std::array<int, 4>& t, int v) const;
    ExactPositions& requireExact(const char* who) const;
    // the device part of isFlipAt: orientations on posf (v at *at)
    bool posfFlip(const std::vector<std::array<int, 4>>& ts, int v, const Point_3f* at);
    // calTetQualities with v at *at
    void qualitiesAt(const std::vector<std::array<int, 4>>& ts, int v, const Point_3f& at,
                     std::vector<double>& q);
};

class VertexSmoother : public LocalOperations {
public:
    using LocalOperations::LocalOperations;

    // src/tetwild/VertexSmoother.h:23-30.  smoothSingle / smoothSurface give
    // the reference's vertex-order (Gauss-Seidel) results, computed level by
    // level: a vertex's level is one more than the highest level of the
    // earlier vertices it shares a tet with, so each level is a set of
    // independent vertices that sees exactly the neighbours' updates the
    // vertex-order loop would show it; every level is one batch.
    void smooth();
    void smoothSingle();
    void smoothSurface();
    // the same loops one vertex per call (the same results, one launch each)
    void smoothSingleSerial();
    void smoothSurfaceSerial();
    // levels of the last smoothSingle / smoothSurface call
    int last_levels = 0;
    bool NewtonsMethod(const std::vector<int>& t_ids,
                       const std::vector<std::array<int, 4>>& new_tets, int v_id, Point_3f& p);
    bool NewtonsUpdate(const std::vector<int>& t_ids, int v_id, double& energy,
                       std::array<double, 3>& J, std::array<double, 9>& H,
                       std::array<double, 3>& X0);
    double getNewEnergy(const std::vector<int>& t_ids);

    // Batched variants: vertices of one colour class (no two share a tet) are
    // smoothed in one persistent launch; classes run one after another, each
    // seeing the positions the previous ones produced.  Same per-vertex rule
    // as smoothSingle / smoothSurface, different (colour-major) order.
    void smoothSingleBatched();
    void smoothSurfaceBatched();

    // one-ring of v in the order the mirrors use (sorted tet ids)
    void oneRing(int v_id, std::vector<int>& t_ids) const;

private:
    struct Problem {
        int v;
        int kind;
    };
    // one vertex of smoothSingle (kind 0) or smoothSurface (kind 1 / 2) from
    // the try-to-round step on; returns the TWG_SMOOTH_* status
    int smoothOne(int v, int kind);
    // smoothSingle / smoothSurface's try-to-round step
    // (src/tetwild/VertexSmoother.cpp:152-161, 272-281)
    void tryRound(int v, const std::vector<std::array<int, 4>>& ring);
    // a ring with an unrounded vertex other than v: its flip tests go to
    // ExactPositions, so the problem runs step by step (twg_newton_step)
    int smoothExact(int v, int kind, const std::vector<int>& ring);
    bool newtonExact(int v, const std::vector<int>& ring, Point_3f& p);
    // rounded-vertex bookkeeping after a problem's outcome
    void afterMove(int v, int status);
    // one colour class of the batched variants; suc counts its successes
    void runClass(const std::vector<int>& cls, bool surface, int& suc);
    // runs the problems (all independent) as one batch; returns statuses and
    // writes moved positions / qualities back into the mesh
    void runBatch(const std::vector<Problem>& probs, std::vector<uint8_t>& status);
    std::vector<std::vector<int>> colourClasses(const std::vector<int>& verts) const;
    // verts (in vertex order) grouped by dependency level
    std::vector<std::vector<int>> orderLevels(const std::vector<int>& verts) const;
    // every live vertex is rounded (no test needs ExactPositions)
    bool allRounded() const;
    // the levels with positions and qualities kept on the device: per level
    // a quality gather, one batch and the write-back, no host round trip
    // between levels; the same updates and counts as runClass per level
    void runLevelsOnDevice(const std::vector<std::vector<int>>& levels, bool surface, int& suc);
    // the surface loop's flag updates (isolated vertices leave the surface,
    // non-boundary points the boundary), in vertex order; returns the
    // vertices smoothSurface smooths
    std::vector<int> surfaceCandidates();
    void finishSingle();
    std::vector<double> last_pout_;  // p_out of the last batch
};

// The checks of EdgeCollapser::collapseAnEdge (src/tetwild/EdgeCollapser.h:
// 23-92, EdgeCollapser.cpp:284-446) up to its "real update": the caller (the
// collapse queue and the topology update) stays on the host.
class EdgeCollapser : public LocalOperations {
public:
    using LocalOperations::LocalOperations;

    // return codes, src/tetwild/EdgeCollapser.h:52-56
    static constexpr int SUCCESS = 0;
    static constexpr int FLIP = 1;
    static constexpr int QUALITY = 2;
    static constexpr int ENVELOP = 3;
    static constexpr int ENVELOP_SUC = 4;

    int energy_type = State::ENERGY_AMIPS;
    bool is_check_quality = true;
    bool is_soft = false;
    double soft_energy = 6;

    // collapseAnEdge's checks for v1 -> v2, including its updates of v1's
    // surface / boundary flags (:298-307).  new_qs receives the new tets'
    // qualities in one-ring order (the removed tets skipped), which the caller
    // stores when the code is SUCCESS or ENVELOP_SUC.
    int checkCollapse(int v1_id, int v2_id, std::vector<TetQuality>* new_qs = nullptr);
    // Many candidates in one launch, each checked against the mesh as it is
    // now (no collapse is applied in between); the flag updates run in
    // candidate order first, as the sequential checks would run them.
    void checkCollapses(const std::vector<std::array<int, 2>>& edges, std::vector<int>& codes,
                        std::vector<std::vector<TetQuality>>* new_qs = nullptr);

    // v1's one-ring in the order the mirrors use (sorted tet ids)
    void oneRing(int v_id, std::vector<int>& t_ids) const;

private:
    // collapseAnEdge's checks in the reference's order on the host, each
    // through the mirrors' device calls: the candidates whose new tets hold
    // an unrounded vertex (their flip test needs ExactPositions)
    int checkCollapseExact(int v1_id, int v2_id, std::vector<TetQuality>* new_qs);
};

// The decision of EdgeRemover::swap for an edge (src/tetwild/EdgeRemover.h,
// EdgeRemover.cpp:97-617): which of removeAnEdge_32 / _44 / _56 succeeds and
// the tets and qualities it would write; the queue and the topology update
// stay with the caller.
class EdgeRemover : public LocalOperations {
public:
    using LocalOperations::LocalOperations;

    int energy_type = State::ENERGY_AMIPS;
    int equal_buget = 100;  // src/tetwild/EdgeRemover.cpp:60

    struct Removal {
        int kind = 0;  // 3, 4 or 5: the operator that succeeds; 0: none
        std::vector<std::array<int, 4>> new_tets;
        std::vector<TetQuality> new_qs;
    };

    // the edge (v1, v2) with its tets old_t_ids (the reference's order);
    // equal_buget is consumed as the reference consumes it
    Removal checkRemoval(int v1_id, int v2_id, const std::vector<int>& old_t_ids);
    // many edges: every configuration evaluated in one launch, the decisions
    // taken in edge order (equal_buget carried through), each against the mesh
    // as it is now
    void checkRemovals(const std::vector<std::array<int, 2>>& edges,
                       const std::vector<std::vector<int>>& rings, std::vector<Removal>& out);
};

}  // namespace tetwild_amd
