/*
 * tetwild_gpu.h -- C-ABI of the MI355X mesh-optimisation inner loop
 * (libtetwild_amd.so).
 *
 * Plain pointers and sizes only.  Unless a function says otherwise, array
 * arguments are DEVICE pointers (HIP global memory) and the call is enqueued
 * on the context's stream and returns before the work finishes; call
 * twg_ctx_sync() (or synchronise the stream) before reading results.  Every
 * function returns TWG_OK (0) or a nonzero status; twg_last_error() gives the
 * message of the calling thread's last failure.  A context is not
 * thread-safe: callers serialise the calls they make on one context.
 *
 * Layouts (all little-endian, naturally aligned):
 *   vertices  three float64 arrays x[nv], y[nv], z[nv]          (SoA)
 *             a tet whose vertex ids fall outside [0, nv) gets NaN outputs
 *             (flip test: flipped) instead of an out-of-bounds read
 *   tets      int32 [n][4] vertex ids, 16-byte aligned          (AoS, 16 B)
 *   gradients float64 planes G[k*n + i], k = 0..2               (SoA)
 *   Hessians  float64 planes H[k*n + i], k = 0..8 row-major     (SoA)
 *   triangles float64 [n][9] = v0xyz v1xyz v2xyz                (AoS)
 *   points    float64 [n][3]                                    (AoS)
 *
 * Reference interfaces replaced (paths relative to the TetWild sources):
 *   see each declaration.
 */
#ifndef TETWILD_GPU_H
#define TETWILD_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TWG_OK 0
#define TWG_EINVAL 1   /* bad argument (null pointer, negative size, ...) */
#define TWG_EDEVICE 2  /* HIP runtime error */
#define TWG_ENOMEM 3   /* device allocation failed */

#define TWG_FP64 0     /* FP64 arithmetic, bit-exact with the CPU restatement */
#define TWG_MIXED 1    /* FP32 bulk with certified bounds, FP64 fallback */

#define TWG_AMIPS_E 0   /* energy */
#define TWG_AMIPS_EG 1  /* energy + gradient w.r.t. vertex 0 */
#define TWG_AMIPS_EGH 2 /* energy + gradient + Hessian w.r.t. vertex 0 */

#define TWG_MAX_ENERGY 1e50 /* State::MAX_ENERGY, src/tetwild/State.h:29 */

typedef struct twg_ctx twg_ctx;
typedef struct twg_bvh twg_bvh;

/* ---- context, memory ------------------------------------------------- */
/* Creates a context on HIP device `device` with its own (default-priority)
 * stream.  No reference counterpart: the reference runs on the host thread. */
int twg_ctx_create(int device, twg_ctx** out);
int twg_ctx_destroy(twg_ctx* ctx);
/* Enqueue later calls on an external HIP stream (hipStream_t; NULL is the HIP
 * null stream).  twg_ctx_stream() returns the stream in use (initially the
 * context's own). */
int twg_ctx_set_stream(twg_ctx* ctx, void* hip_stream);
void* twg_ctx_stream(twg_ctx* ctx);
int twg_ctx_sync(twg_ctx* ctx);
const char* twg_last_error(void);
/* Device memory helpers for hosts that do not link HIP themselves.  The
 * copies are synchronous with respect to the host. */
int twg_malloc(twg_ctx* ctx, size_t bytes, void** dptr);
int twg_free(twg_ctx* ctx, void* dptr);
int twg_memcpy_h2d(twg_ctx* ctx, void* dst, const void* src, size_t bytes);
int twg_memcpy_d2h(twg_ctx* ctx, void* dst, const void* src, size_t bytes);

/* ---- (a) AMIPS energy, gradient, Hessian ------------------------------ */
/* Per tet i (vertex ids tets[i][0..3]): conformal AMIPS energy E[i], its
 * gradient G and Hessian H w.r.t. vertex tets[i][0].
 * Replaces LocalOperations::comformalAMIPSEnergy_new / Jacobian_new /
 * Hessian_new (src/tetwild/LocalOperations.h:109-111,
 * src/tetwild/LocalOperations.cpp:28-291) and energy_ispc
 * (src/ispc/energy.ispc:7-65), batched.  mode: TWG_AMIPS_E / _EG / _EGH; G and
 * H may be NULL when the mode does not write them.  FP64, bit-exact with the
 * CPU restatement; within ~1e-12 relative of the reference's code-generated
 * expressions on well-shaped tets. */
int twg_amips_eval(twg_ctx* ctx, const int32_t* tets, int64_t n, const double* x,
                   const double* y, const double* z, int64_t nv, int mode, double* E, double* G,
                   double* H);

/* Mixed-precision variant: FP32 arithmetic from FP64 coordinates, FP32
 * outputs for G and H, and an energy E[i] (float64) that is either the FP32
 * value, certified to relative error <= tol by a forward error bound, or, where
 * the bound exceeds tol (ill-conditioned tets), the FP64 value.  fallback[i]
 * (may be NULL) is 1 where the FP64 value was used.  E is finite-or-inf exactly
 * where the FP64 energy is. */
int twg_amips_eval_mixed(twg_ctx* ctx, const int32_t* tets, int64_t n, const double* x,
                         const double* y, const double* z, int64_t nv, int mode, double tol,
                         double* E, float* G, float* H, uint8_t* fallback);

/* calTetQualities (src/tetwild/LocalOperations.h:69,
 * src/tetwild/LocalOperations.cpp:695-773 with calTetQuality_AMIPS :862-884):
 * slim_energy[i] = MAX_ENERGY if the exact orientation of the tet is not
 * POSITIVE, else the AMIPS energy with inf / nan / <= 0 mapped to MAX_ENERGY.
 * precision TWG_MIXED computes energies as twg_amips_eval_mixed with tol. */
int twg_tet_qualities(twg_ctx* ctx, const int32_t* tets, int64_t n, const double* x,
                      const double* y, const double* z, int64_t nv, int precision, double tol,
                      double* slim_energy, uint8_t* fallback);

/* isTetFlip on rounded vertices (src/tetwild/LocalOperations.cpp:535-553):
 * flipped[i] = 1 unless the exact orientation of tet i is POSITIVE. */
int twg_tet_flip(twg_ctx* ctx, const int32_t* tets, int64_t n, const double* x,
                 const double* y, const double* z, int64_t nv, uint8_t* flipped);

#ifdef __cplusplus
}
#endif
#endif
