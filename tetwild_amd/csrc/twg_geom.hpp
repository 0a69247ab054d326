// twg_geom.hpp -- geometry of the envelope test: vec3 helpers with geogram's
// evaluation order, Eberly's point-triangle distance, triangle sampling and
// point-box distances.  FP64 expressions follow the same operation order as
// the host restatement, so their results are bit-identical.
#pragma once

#include "twg_math.hpp"

namespace twg {

template <typename R>
struct V3 {
    R x, y, z;
};

template <typename R>
TWG_HD V3<R> mk3(R x, R y, R z)
{
    V3<R> r;
    r.x = x;
    r.y = y;
    r.z = z;
    return r;
}
template <typename R> TWG_HD V3<R> operator-(V3<R> a, V3<R> b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
template <typename R> TWG_HD V3<R> operator+(V3<R> a, V3<R> b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
template <typename R> TWG_HD V3<R> mulv(V3<R> a, R s) { return mk3(a.x * s, a.y * s, a.z * s); }
template <typename R> TWG_HD V3<R> divv(V3<R> a, R s) { return mk3(a.x / s, a.y / s, a.z / s); }
template <typename R> TWG_HD R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename R> TWG_HD R len2(V3<R> a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
TWG_HD double len(V3<double> a) { return __builtin_sqrt(len2(a)); }
TWG_HD double dist(V3<double> a, V3<double> b) { return len(a - b); }
template <typename R>
TWG_HD V3<R> cross(V3<R> a, V3<R> b)
{
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// geogram normalize: v * (1 / |v|) unless |v| <= 1e-30
TWG_HD V3<double> normalize(V3<double> v)
{
    double s = len(v);
    if (s > 1e-30) s = 1.0 / s;
    return mulv(v, s);
}

template <typename R>
TWG_HD R absr(R v) { return v < R(0) ? -v : v; }

// ---- point - segment / point - triangle squared distance ---------------
template <typename R>
TWG_HD R point_seg_d2(V3<R> p, V3<R> a, V3<R> b, V3<R>& cp)
{
    R l2 = len2(a - b);
    R t = dot(p - a, b - a);
    if (t <= R(0) || l2 == R(0)) {
        cp = a;
        return len2(p - a);
    }
    if (t > l2) {
        cp = b;
        return len2(p - b);
    }
    R l1 = t / l2;
    R l0 = R(1) - l1;
    cp = mulv(a, l0) + mulv(b, l1);
    return len2(p - cp);
}

// Eberly, "Distance Between Point and Triangle in 3D" (region method); the
// method behind geogram's point_triangle_squared_distance.
template <typename R, bool WANT_CP>
TWG_HD R point_tri_d2(V3<R> p, V3<R> V0, V3<R> V1, V3<R> V2, V3<R>& cp)
{
    V3<R> diff = V0 - p;
    V3<R> e0 = V1 - V0;
    V3<R> e1 = V2 - V0;
    R a00 = len2(e0);
    R a01 = dot(e0, e1);
    R a11 = len2(e1);
    R b0 = dot(diff, e0);
    R b1 = dot(diff, e1);
    R c = len2(diff);
    R det = absr(a00 * a11 - a01 * a01);
    R s = a01 * b1 - a11 * b0;
    R t = a01 * b0 - a00 * b1;
    R d2;
    const R one = R(1), zero = R(0), two = R(2);

    if (det < R(1e-30)) {
        V3<R> q;
        R best = point_seg_d2(p, V0, V1, cp);
        R cur = point_seg_d2(p, V0, V2, q);
        if (cur < best) { best = cur; cp = q; }
        cur = point_seg_d2(p, V1, V2, q);
        if (cur < best) { best = cur; cp = q; }
        return best;
    }
    if (s + t <= det) {
        if (s < zero) {
            if (t < zero) {
                if (b0 < zero) {
                    t = zero;
                    if (-b0 >= a00) { s = one; d2 = a00 + two * b0 + c; }
                    else { s = -b0 / a00; d2 = b0 * s + c; }
                } else {
                    s = zero;
                    if (b1 >= zero) { t = zero; d2 = c; }
                    else if (-b1 >= a11) { t = one; d2 = a11 + two * b1 + c; }
                    else { t = -b1 / a11; d2 = b1 * t + c; }
                }
            } else {
                s = zero;
                if (b1 >= zero) { t = zero; d2 = c; }
                else if (-b1 >= a11) { t = one; d2 = a11 + two * b1 + c; }
                else { t = -b1 / a11; d2 = b1 * t + c; }
            }
        } else if (t < zero) {
            t = zero;
            if (b0 >= zero) { s = zero; d2 = c; }
            else if (-b0 >= a00) { s = one; d2 = a00 + two * b0 + c; }
            else { s = -b0 / a00; d2 = b0 * s + c; }
        } else {
            R inv = one / det;
            s *= inv;
            t *= inv;
            d2 = s * (a00 * s + a01 * t + two * b0) + t * (a01 * s + a11 * t + two * b1) + c;
        }
    } else {
        if (s < zero) {
            R tmp0 = a01 + b0;
            R tmp1 = a11 + b1;
            if (tmp1 > tmp0) {
                R numer = tmp1 - tmp0;
                R denom = a00 - two * a01 + a11;
                if (numer >= denom) { s = one; t = zero; d2 = a00 + two * b0 + c; }
                else {
                    s = numer / denom;
                    t = one - s;
                    d2 = s * (a00 * s + a01 * t + two * b0) + t * (a01 * s + a11 * t + two * b1) + c;
                }
            } else {
                s = zero;
                if (tmp1 <= zero) { t = one; d2 = a11 + two * b1 + c; }
                else if (b1 >= zero) { t = zero; d2 = c; }
                else { t = -b1 / a11; d2 = b1 * t + c; }
            }
        } else if (t < zero) {
            R tmp0 = a01 + b1;
            R tmp1 = a00 + b0;
            if (tmp1 > tmp0) {
                R numer = tmp1 - tmp0;
                R denom = a00 - two * a01 + a11;
                if (numer >= denom) { t = one; s = zero; d2 = a11 + two * b1 + c; }
                else {
                    t = numer / denom;
                    s = one - t;
                    d2 = s * (a00 * s + a01 * t + two * b0) + t * (a01 * s + a11 * t + two * b1) + c;
                }
            } else {
                t = zero;
                if (tmp1 <= zero) { s = one; d2 = a00 + two * b0 + c; }
                else if (b0 >= zero) { s = zero; d2 = c; }
                else { s = -b0 / a00; d2 = b0 * s + c; }
            }
        } else {
            R numer = a11 + b1 - a01 - b0;
            if (numer <= zero) { s = zero; t = one; d2 = a11 + two * b1 + c; }
            else {
                R denom = a00 - two * a01 + a11;
                if (numer >= denom) { s = one; t = zero; d2 = a00 + two * b0 + c; }
                else {
                    s = numer / denom;
                    t = one - s;
                    d2 = s * (a00 * s + a01 * t + two * b0) + t * (a01 * s + a11 * t + two * b1) + c;
                }
            }
        }
    }
    if (d2 < zero) d2 = zero;
    if (WANT_CP) cp = (V0 + mulv(e0, s)) + mulv(e1, t);
    return d2;
}

// x0, x1 or x2 by i, component-wise selects on values
__device__ __forceinline__ V3<double> sel3(int i, V3<double> x0, V3<double> x1, V3<double> x2)
{
    return mk3(i == 0 ? x0.x : (i == 1 ? x1.x : x2.x), i == 0 ? x0.y : (i == 1 ? x1.y : x2.y),
               i == 0 ? x0.z : (i == 1 ? x1.z : x2.z));
}

// ---- sampleTriangle (src/tetwild/Common.cpp:143-228) -------------------
// A resumable generator: next(p) yields the samples one at a time in the
// reference's generation order (longest edge, v1, rows parallel to it, v2,
// then the two other edges), so a caller has a single consumption site.
struct SampleGen {
    // State kept between samples: the rotated triangle, the spacing, the
    // current edge's sample count, the current row's start and the counters.
    // The unit directions and the angles of the rows are recomputed where they
    // are used, with the same expressions (so every sample is bit-identical):
    // kept, they would hold registers across the envelope traversal that runs
    // between two samples (the step direction alone, kept, made the mixed
    // face kernel 6 % slower).  skip(), which runs no traversal, computes the
    // rows' angles once per call.
    V3<double> v0, v1, v2, rowv;
    double d, N;
    int phase, k, m, M, N1;

    // what every row header of the triangle uses
    struct RowGeom {
        V3<double> n02, n12;
        double sin_v0, tan_v0, sin_v1;
    };

    __device__ __forceinline__ static double sqrt3_2() { return __builtin_sqrt(3.0) / 2; }

    // the triangle arrives as three values (not an array): rotating an array
    // by a run-time index would place it in scratch memory
    __device__ __forceinline__ void init(V3<double> a, V3<double> b, V3<double> c, double d_)
    {
        d = d_;
        // squared edge lengths |v_i - v_(i+1)|^2
        const double l0 = len2(a - b), l1 = len2(b - c), l2 = len2(c - a);
        int max_i = 0;
        double lsm = l0;
        if (!(l1 < lsm)) { max_i = 1; lsm = l1; }  // last largest, as std::minmax_element
        if (!(l2 < lsm)) { max_i = 2; lsm = l2; }
        N = __builtin_sqrt(lsm) / d;
        k = 0;
        if (N <= 1) {
            v0 = a;
            v1 = b;
            v2 = c;
            phase = 9;
            return;
        }
        if (N == (int)N) N -= 1;
        v0 = sel3(max_i, a, b, c);
        v1 = sel3(max_i, b, c, a);
        v2 = sel3(max_i, c, a, b);
        V3<double> e01 = v1 - v0;
        double h = dist(divv(mulv(e01, dot(v2 - v0, e01)), lsm) + v0, v2);
        M = (int)(h / (sqrt3_2() * d));
        phase = 0;
    }

    __device__ __forceinline__ RowGeom row_geom() const
    {
        RowGeom g;
        g.n02 = normalize(v2 - v0);
        g.n12 = normalize(v2 - v1);
        g.sin_v0 = len(cross(v2 - v0, v1 - v0)) / (dist(v0, v2) * dist(v0, v1));
        g.tan_v0 = len(cross(v2 - v0, v1 - v0)) / dot(v2 - v0, v1 - v0);
        g.sin_v1 = len(cross(v2 - v1, v0 - v1)) / (dist(v1, v2) * dist(v0, v1));
        return g;
    }

    // phase 2: the header of row m -- its start and sample count (phase 3),
    // or the end of the rows (phase 4)
    __device__ __forceinline__ void row_header(const RowGeom& g)
    {
        const double s32 = sqrt3_2();
        int n = (int)(s32 / g.tan_v0 * m + 0.5);
        int n1 = (int)(s32 / g.tan_v0 * m);
        if (m % 2 == 0 && n == n1) n += 1;
        V3<double> v0_m = v0 + mulv(g.n02, m * s32 * d / g.sin_v0);
        V3<double> v1_m = v1 + mulv(g.n12, m * s32 * d / g.sin_v1);
        if (dist(v0_m, v1_m) <= d) {
            phase = 4;
            return;
        }
        double delta_d = ((n + (m % 2) / 2.0) - m * s32 / g.tan_v0) * d;
        rowv = v0_m + mulv(normalize(v1 - v0), delta_d);
        N1 = (int)(dist(rowv, v1_m) / d);
        k = 0;
        phase = 3;
    }
    __device__ __forceinline__ void row_header()
    {
        if (m > M) {
            phase = 4;
            return;
        }
        row_header(row_geom());
    }

    // after v2: the edge v1 -> v2 (|v1 - v2|^2 is the edge's squared length)
    __device__ __forceinline__ void start_edge12()
    {
        N = __builtin_sqrt(len2(v1 - v2)) / d;
        if (N > 1) {
            if (N == (int)N) N -= 1;
            k = 1;
            phase = 5;
        } else {
            phase = 6;
        }
    }

    // the edge v2 -> v0; false when it has no interior samples (the end)
    __device__ __forceinline__ bool start_edge20()
    {
        N = __builtin_sqrt(len2(v2 - v0)) / d;
        if (!(N > 1)) return false;
        if (N == (int)N) N -= 1;
        k = 1;
        phase = 7;
        return true;
    }

    // Moves past the next s samples without producing them (a row is passed
    // over from its header alone) and returns how many it passed (fewer than
    // s at the end of the sequence).  next() then yields the sample after
    // them, bit-identical to stepping through one at a time.
    __device__ __forceinline__ int64_t skip(int64_t s)
    {
        int64_t done = 0;
        RowGeom g;
        bool have_g = false;
        for (;;) {
            if (done >= s) return done;
            const int64_t want = s - done;
            switch (phase) {
            case 9:
                if (k >= 3) return done;
                {
                    const int64_t take = want < 3 - k ? want : 3 - k;
                    k += (int)take;
                    done += take;
                }
                break;
            case 0:
                if (k <= N) {
                    const int64_t avail = (int64_t)N - k + 1;
                    const int64_t take = want < avail ? want : avail;
                    k += (int)take;
                    done += take;
                } else {
                    phase = 1;  // v1
                    done += 1;
                }
                break;
            case 1:
                if (M < 1) {
                    phase = 8;  // v2, the last sample
                    done += 1;
                } else {
                    m = 1;
                    phase = 2;
                }
                break;
            case 2:
                if (m > M) {
                    phase = 4;
                } else {
                    if (!have_g) {
                        g = row_geom();
                        have_g = true;
                    }
                    row_header(g);
                }
                break;
            case 3:
                if (k <= N1) {
                    const int64_t avail = (int64_t)N1 - k + 1;
                    const int64_t take = want < avail ? want : avail;
                    k += (int)take;
                    done += take;
                } else {
                    m++;
                    phase = 2;
                }
                break;
            case 4:  // v2
                start_edge12();
                done += 1;
                break;
            case 5:
                if (k <= N) {
                    const int64_t avail = (int64_t)N - k + 1;
                    const int64_t take = want < avail ? want : avail;
                    k += (int)take;
                    done += take;
                } else {
                    phase = 6;
                }
                break;
            case 6:
                if (!start_edge20()) {
                    phase = 8;
                    return done;
                }
                break;
            case 7:
                if (k <= N) {
                    const int64_t avail = (int64_t)N - k + 1;
                    const int64_t take = want < avail ? want : avail;
                    k += (int)take;
                    done += take;
                } else {
                    phase = 8;
                    return done;
                }
                break;
            default:
                return done;
            }
        }
    }

    // Writes what SampleGenC needs after init(): the (rotated) triangle, the
    // edge directions, the rows' angles and the edge counts, each computed
    // with the expression next() and skip() use.
    template <class P>
    __device__ __forceinline__ void store(P rec) const;

    // Writes the triangle's segment record and header after init() (the
    // segmented layout below); row0: the first row-table entry of its rows.
    template <class P, class I>
    __device__ __forceinline__ void store_seg(P rec, I hdr, int row0) const;

    // the samples of an edge of L sampling distances past its first vertex
    // (start_edge12 / start_edge20: k = 1 .. N)
    __device__ __forceinline__ static int edge_count(double L)
    {
        if (!(L > 1)) return 0;
        if (L == (int)L) L -= 1;
        return (int)L;
    }

    __device__ __forceinline__ bool next(V3<double>& p)
    {
        for (;;) {
            switch (phase) {
            case 9:  // short triangle: its three vertices
                if (k < 3) {
                    p = sel3(k, v0, v1, v2);
                    k++;
                    return true;
                }
                return false;
            case 0:  // longest edge
                if (k <= N) {
                    p = v0 + mulv(mulv(normalize(v1 - v0), d), (double)k);
                    k++;
                    return true;
                }
                phase = 1;
                p = v1;
                return true;
            case 1:  // after v1: rows, or straight to v2
                if (M < 1) {
                    phase = 8;
                    p = v2;
                    return true;
                }
                m = 1;
                phase = 2;
                break;
            case 2:  // row header
                row_header();
                break;
            case 3:  // row samples
                if (k <= N1) {
                    p = rowv + mulv(mulv(normalize(v1 - v0), (double)k), d);
                    k++;
                    return true;
                }
                m++;
                phase = 2;
                break;
            case 4:  // v2, then the edge v1 -> v2
                p = v2;
                start_edge12();
                return true;
            case 5:
                if (k <= N) {
                    p = v1 + mulv(mulv(normalize(v2 - v1), d), (double)k);
                    k++;
                    return true;
                }
                phase = 6;
                break;
            case 6:  // the edge v2 -> v0
                if (!start_edge20()) return false;
                break;
            case 7:
                if (k <= N) {
                    p = v2 + mulv(mulv(normalize(v0 - v2), d), (double)k);
                    k++;
                    return true;
                }
                return false;
            case 8:  // done after v2 (no rows)
            default:
                return false;
            }
        }
    }
};

// Layout of a SampleGen::store record (doubles).
enum : int {
    kFrV0 = 0, kFrV1 = 3, kFrV2 = 6,        // the rotated triangle
    kFrN01 = 9, kFrN02 = 12, kFrN12 = 15,   // normalize(v1 - v0), (v2 - v0), (v2 - v1)
    kFrN20 = 18,                            // normalize(v0 - v2)
    kFrSin0 = 21, kFrTan0 = 22, kFrSin1 = 23,
    kFrN = 24,                              // the longest edge's count after init
    kFrL12 = 25, kFrL20 = 26,               // |v1 - v2| / d, |v2 - v0| / d
    kFrM = 27, kFrPhase = 28,
    kFaceRec = 29
};

template <class P>
__device__ __forceinline__ void SampleGen::store(P r) const
{
    const V3<double> vs[3] = {v0, v1, v2};
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        r[kFrV0 + 3 * j] = vs[j].x;
        r[kFrV0 + 3 * j + 1] = vs[j].y;
        r[kFrV0 + 3 * j + 2] = vs[j].z;
    }
    r[kFrPhase] = (double)phase;
    r[kFrN] = N;
    r[kFrM] = phase == 9 ? 0.0 : (double)M;
    if (phase == 9) return;
    const RowGeom g = row_geom();
    const V3<double> n01 = normalize(v1 - v0), n20 = normalize(v0 - v2);
    const V3<double> ds[4] = {n01, g.n02, g.n12, n20};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        r[kFrN01 + 3 * j] = ds[j].x;
        r[kFrN01 + 3 * j + 1] = ds[j].y;
        r[kFrN01 + 3 * j + 2] = ds[j].z;
    }
    r[kFrSin0] = g.sin_v0;
    r[kFrTan0] = g.tan_v0;
    r[kFrSin1] = g.sin_v1;
    r[kFrL12] = __builtin_sqrt(len2(v1 - v2)) / d;
    r[kFrL20] = __builtin_sqrt(len2(v2 - v0)) / d;
}

// SampleGen over a stored record (P: a pointer to its doubles, e.g. in LDS):
// the same samples bit for bit, the same skip(), with the directions, the
// rows' angles and the edge lengths read instead of recomputed.  The state
// between samples is the record pointer, the row start and the counters.
template <class P>
struct SampleGenC {
    P r;
    V3<double> rowv;
    double d, N;
    int phase, k, m, M, N1;

    __device__ __forceinline__ V3<double> at(int o) const { return mk3(r[o], r[o + 1], r[o + 2]); }

    __device__ __forceinline__ void load(P rec, double d_)
    {
        r = rec;
        d = d_;
        phase = (int)r[kFrPhase];
        N = r[kFrN];
        M = (int)r[kFrM];
        k = 0;
    }

    __device__ __forceinline__ void row_header()
    {
        if (m > M) {
            phase = 4;
            return;
        }
        const double s32 = SampleGen::sqrt3_2();
        const double tan_v0 = r[kFrTan0];
        int n = (int)(s32 / tan_v0 * m + 0.5);
        int n1 = (int)(s32 / tan_v0 * m);
        if (m % 2 == 0 && n == n1) n += 1;
        const V3<double> v0 = at(kFrV0), v1 = at(kFrV1);
        V3<double> v0_m = v0 + mulv(at(kFrN02), m * s32 * d / r[kFrSin0]);
        V3<double> v1_m = v1 + mulv(at(kFrN12), m * s32 * d / r[kFrSin1]);
        if (dist(v0_m, v1_m) <= d) {
            phase = 4;
            return;
        }
        double delta_d = ((n + (m % 2) / 2.0) - m * s32 / tan_v0) * d;
        rowv = v0_m + mulv(at(kFrN01), delta_d);
        N1 = (int)(dist(rowv, v1_m) / d);
        k = 0;
        phase = 3;
    }

    __device__ __forceinline__ void start_edge12()
    {
        N = r[kFrL12];
        if (N > 1) {
            if (N == (int)N) N -= 1;
            k = 1;
            phase = 5;
        } else {
            phase = 6;
        }
    }

    __device__ __forceinline__ bool start_edge20()
    {
        N = r[kFrL20];
        if (!(N > 1)) return false;
        if (N == (int)N) N -= 1;
        k = 1;
        phase = 7;
        return true;
    }

    // SampleGen::skip
    __device__ __forceinline__ int64_t skip(int64_t s)
    {
        int64_t done = 0;
        for (;;) {
            if (done >= s) return done;
            const int64_t want = s - done;
            switch (phase) {
            case 9:
                if (k >= 3) return done;
                {
                    const int64_t take = want < 3 - k ? want : 3 - k;
                    k += (int)take;
                    done += take;
                }
                break;
            case 0:
                if (k <= N) {
                    const int64_t avail = (int64_t)N - k + 1;
                    const int64_t take = want < avail ? want : avail;
                    k += (int)take;
                    done += take;
                } else {
                    phase = 1;
                    done += 1;
                }
                break;
            case 1:
                if (M < 1) {
                    phase = 8;
                    done += 1;
                } else {
                    m = 1;
                    phase = 2;
                }
                break;
            case 2:
                row_header();
                break;
            case 3:
                if (k <= N1) {
                    const int64_t avail = (int64_t)N1 - k + 1;
                    const int64_t take = want < avail ? want : avail;
                    k += (int)take;
                    done += take;
                } else {
                    m++;
                    phase = 2;
                }
                break;
            case 4:
                start_edge12();
                done += 1;
                break;
            case 5:
                if (k <= N) {
                    const int64_t avail = (int64_t)N - k + 1;
                    const int64_t take = want < avail ? want : avail;
                    k += (int)take;
                    done += take;
                } else {
                    phase = 6;
                }
                break;
            case 6:
                if (!start_edge20()) {
                    phase = 8;
                    return done;
                }
                break;
            case 7:
                if (k <= N) {
                    const int64_t avail = (int64_t)N - k + 1;
                    const int64_t take = want < avail ? want : avail;
                    k += (int)take;
                    done += take;
                } else {
                    phase = 8;
                    return done;
                }
                break;
            default:
                return done;
            }
        }
    }

    // SampleGen::next
    __device__ __forceinline__ bool next(V3<double>& p)
    {
        for (;;) {
            switch (phase) {
            case 9:
                if (k < 3) {
                    p = at(kFrV0 + 3 * k);
                    k++;
                    return true;
                }
                return false;
            case 0:
                if (k <= N) {
                    p = at(kFrV0) + mulv(mulv(at(kFrN01), d), (double)k);
                    k++;
                    return true;
                }
                phase = 1;
                p = at(kFrV1);
                return true;
            case 1:
                if (M < 1) {
                    phase = 8;
                    p = at(kFrV2);
                    return true;
                }
                m = 1;
                phase = 2;
                break;
            case 2:
                row_header();
                break;
            case 3:
                if (k <= N1) {
                    p = rowv + mulv(mulv(at(kFrN01), (double)k), d);
                    k++;
                    return true;
                }
                m++;
                phase = 2;
                break;
            case 4:
                p = at(kFrV2);
                start_edge12();
                return true;
            case 5:
                if (k <= N) {
                    p = at(kFrV1) + mulv(mulv(at(kFrN12), d), (double)k);
                    k++;
                    return true;
                }
                phase = 6;
                break;
            case 6:
                if (!start_edge20()) return false;
                break;
            case 7:
                if (k <= N) {
                    p = at(kFrV2) + mulv(mulv(at(kFrN20), d), (double)k);
                    k++;
                    return true;
                }
                return false;
            case 8:
            default:
                return false;
            }
        }
    }
};

// ---- sampleTriangle as segments ------------------------------------------
// SampleGen's sequence laid out for a group of lanes that sample several
// triangles at once.  A triangle's samples are runs along lines: the longest
// edge v0 -> v1, the vertex v1, the rows, v2, the edges v1 -> v2 and
// v2 -> v0, each sample the run's start plus a multiple of its step with the
// expression SampleGen uses for that run.  A triangle keeps its rotated
// vertices and unit directions (kSeg* doubles) and a header of counts (kSh*
// ints); its rows sit in a row table, one entry (the row's start, its sample
// count) per row, each filled on its own by seg_row (a row does not depend on
// the rows before it).  Segment s of a triangle with R rows:
//   short triangle (N <= 1): s = 0, 1, 2 its vertices;
//   no rows (M < 1):         0 the edge v0 -> v1 (C01 samples), 1 v1, 2 v2;
//   otherwise:               0 the edge, 1 v1, 2 .. R + 1 the rows, R + 2 v2,
//                            R + 3 the edge v1 -> v2 (C12), R + 4 v2 -> v0 (C20).
// R is the rows before the first that SampleGen::row_header ends at (the
// header's kShRows, M until the table is read).
enum : int {
    kSegV0 = 0, kSegV1 = 3, kSegV2 = 6,       // the rotated triangle
    kSegN01 = 9, kSegN12 = 12, kSegN20 = 15,  // normalize(v1 - v0), (v2 - v1), (v0 - v2)
    kSegRec = 18
};
enum : int {
    kShKind = 0, kShC01 = 1, kShRows = 2, kShRow0 = 3, kShC12 = 4, kShC20 = 5, kShM = 6,
    kShSum = 7,  // the rows' samples
    kSegHdr = 8
};
enum : int { kSegShort = 0, kSegNoRows = 1, kSegRows = 2 };
constexpr int kSegRow = 4;  // doubles per row-table entry: the start, the count

template <class P>
__device__ __forceinline__ V3<double> seg_at(P r, int o)
{
    return mk3(r[o], r[o + 1], r[o + 2]);
}

template <class P, class I>
__device__ __forceinline__ void SampleGen::store_seg(P r, I h, int row0) const
{
    const V3<double> vs[3] = {v0, v1, v2};
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        r[kSegV0 + 3 * j] = vs[j].x;
        r[kSegV0 + 3 * j + 1] = vs[j].y;
        r[kSegV0 + 3 * j + 2] = vs[j].z;
    }
    h[kShRow0] = row0;
    h[kShSum] = 0;
    if (phase == 9) {
        h[kShKind] = kSegShort;
        h[kShC01] = h[kShRows] = h[kShC12] = h[kShC20] = h[kShM] = 0;
        return;
    }
    const V3<double> ds[3] = {normalize(v1 - v0), normalize(v2 - v1), normalize(v0 - v2)};
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        r[kSegN01 + 3 * j] = ds[j].x;
        r[kSegN01 + 3 * j + 1] = ds[j].y;
        r[kSegN01 + 3 * j + 2] = ds[j].z;
    }
    h[kShKind] = M < 1 ? kSegNoRows : kSegRows;
    h[kShC01] = (int)((int64_t)N + 1);
    h[kShM] = M;
    h[kShRows] = M < 1 ? 0 : M;
    h[kShC12] = edge_count(__builtin_sqrt(len2(v1 - v2)) / d);
    h[kShC20] = edge_count(__builtin_sqrt(len2(v2 - v0)) / d);
}

// Row m (1 <= m <= M) of the triangle whose segment record is r: writes its
// start and returns its count (N1 + 1), or returns 0 when the rows end at m.
template <class P>
__device__ __forceinline__ int seg_row(P r, double d, int m, V3<double>& start)
{
    SampleGen g;
    g.v0 = seg_at(r, kSegV0);
    g.v1 = seg_at(r, kSegV1);
    g.v2 = seg_at(r, kSegV2);
    g.d = d;
    g.m = m;
    g.row_header(g.row_geom());
    if (g.phase != 3) return 0;
    start = g.rowv;
    return g.N1 + 1;
}

// A triangle's header in registers, and its segments over the record r and
// the row table rt (kSegRow doubles per entry).
struct SegFace {
    int kind, c01, rows, row0, c12, c20;

    template <class I>
    __device__ __forceinline__ void load(I h)
    {
        kind = h[kShKind];
        c01 = h[kShC01];
        rows = h[kShRows];
        row0 = h[kShRow0];
        c12 = h[kShC12];
        c20 = h[kShC20];
    }
    __device__ __forceinline__ int nseg() const { return kind == kSegRows ? rows + 5 : 3; }
    // the samples of the triangle, given its rows' samples
    __device__ __forceinline__ int64_t total(int64_t row_sum) const
    {
        return kind == kSegShort ? 3 : kind == kSegNoRows ? (int64_t)c01 + 2 : (int64_t)c01 + 2 + row_sum + c12 + c20;
    }
    template <class RP>
    __device__ __forceinline__ int count(RP rt, int s) const
    {
        if (kind == kSegShort || s == 1) return 1;
        if (s == 0) return c01;
        if (kind == kSegNoRows) return 1;
        const int j = s - 2;
        if (j < rows) return (int)rt[kSegRow * (row0 + j) + 3];
        return j == rows ? 1 : (j == rows + 1 ? c12 : c20);
    }
    // sample k of segment s: a vertex, start + (dir d) k on an edge (k from 1
    // on the edges from v1 and v2), start + (dir k) d on a row -- SampleGen's
    // expressions
    template <class P, class RP>
    __device__ __forceinline__ V3<double> sample(P r, RP rt, int s, int k, double d) const
    {
        int vo = -1, bo = kSegV0, dofs = kSegN01, kk = k;
        bool row = false;
        if (kind == kSegShort) {
            vo = kSegV0 + 3 * s;
        } else if (s == 1) {
            vo = kSegV1;
        } else if (s != 0) {
            const int j = s - 2;
            if (kind == kSegNoRows || j == rows) {
                vo = kSegV2;
            } else if (j < rows) {
                row = true;
            } else {
                const bool e12 = j == rows + 1;
                bo = e12 ? kSegV1 : kSegV2;
                dofs = e12 ? kSegN12 : kSegN20;
                kk = k + 1;
            }
        }
        if (vo >= 0) return seg_at(r, vo);
        const V3<double> dir = seg_at(r, dofs);
        if (row) return seg_at(rt, kSegRow * (row0 + s - 2)) + mulv(mulv(dir, (double)kk), d);
        return seg_at(r, bo) + mulv(mulv(dir, d), (double)kk);
    }
    // (s, k) -> the sample at offset o of the triangle (o below its total;
    // past it, a place in the last segment at or beyond its end)
    template <class RP>
    __device__ __forceinline__ void seek(RP rt, int64_t o, int& s, int& k) const
    {
        const int ns = nseg();
        s = 0;
        while (s + 1 < ns) {
            const int c = count(rt, s);
            if (o < c) break;
            o -= c;
            ++s;
        }
        k = o < INT32_MAX ? (int)o : INT32_MAX;
    }
};

// Calls emit(p) for every sample in order; emit returns true to stop early.
// Returns true iff stopped early.
template <class Emit>
__device__ __forceinline__ bool for_each_sample(V3<double> a, V3<double> b, V3<double> c, double d,
                                                Emit&& emit)
{
    SampleGen g;
    g.init(a, b, c, d);
    V3<double> p;
    while (g.next(p))
        if (emit(p)) return true;
    return false;
}

// ---- boxes ----------------------------------------------------------------
// Unsigned squared distance from p to an axis-aligned box (0 inside).
template <typename R>
TWG_HD R box_d2(V3<R> p, const R* lo, const R* hi)
{
    R r = R(0), d;
    d = p.x < lo[0] ? p.x - lo[0] : (p.x > hi[0] ? p.x - hi[0] : R(0));
    r += d * d;
    d = p.y < lo[1] ? p.y - lo[1] : (p.y > hi[1] ? p.y - hi[1] : R(0));
    r += d * d;
    d = p.z < lo[2] ? p.z - lo[2] : (p.z > hi[2] ? p.z - hi[2] : R(0));
    r += d * d;
    return r;
}

// point_box_signed_squared_distance, src/tetwild/geogram/mesh_AABB.cpp:201-220
TWG_HD double box_signed_d2(V3<double> p, const double* lo, const double* hi)
{
    double pc[3] = {p.x, p.y, p.z};
    bool inside = true;
    double r = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (pc[c] < lo[c]) {
            inside = false;
            double d = pc[c] - lo[c];
            r += d * d;
        } else if (pc[c] > hi[c]) {
            inside = false;
            double d = pc[c] - hi[c];
            r += d * d;
        }
    }
    if (inside) {
        double d = pc[0] - lo[0];
        double in = d * d;
        d = pc[0] - hi[0];
        if (d * d < in) in = d * d;
#pragma unroll
        for (int c = 1; c < 3; ++c) {
            d = pc[c] - lo[c];
            if (d * d < in) in = d * d;
            d = pc[c] - hi[c];
            if (d * d < in) in = d * d;
        }
        r = -in;
    }
    return r;
}

// point_box_center_squared_distance, mesh_AABB.cpp:222-238
TWG_HD double box_center_d2(V3<double> p, const double* lo, const double* hi)
{
    double pc[3] = {p.x, p.y, p.z};
    double r = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double d = pc[c] - 0.5 * (lo[c] + hi[c]);
        r += d * d;
    }
    return r;
}

}  // namespace twg
