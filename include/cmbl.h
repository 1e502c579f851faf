/*
 * cmbl.h -- C ABI of the MI355X-native LenseFlow library (libcmbl.so).
 *
 * The reference (CMBLensing.py, a JAX package) has no FFI: its boundary is the
 * Python call surface.  Each entry point below replaces the compute behind one
 * reference function; the Python package cmblensing.py_amd binds them with
 * ctypes and keeps the reference's names, argument order and defaults.
 *
 * Conventions
 *   - Pointers named in/out/x/X are DEVICE pointers owned by the caller; the
 *     library never frees caller memory.  Plans and phi handles are owned by
 *     the library and freed explicitly.
 *   - MAP arrays are row-major [batch][comp][N][N] real; FOURIER arrays are
 *     [batch][comp][N][N/2+1] interleaved complex (numpy rfft2 layout, axis 0
 *     carries KX, axis 1 the rfft axis KY).  dtype is CMBL_F32 or CMBL_F64 as
 *     fixed by the plan (phi input is always float64).
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *     Work is stream-ordered; calls return after enqueueing unless stated.
 *   - Every call returns 0 on success or a nonzero status; cmbl_last_error()
 *     gives a message (thread-local).  No C++ exception crosses the ABI.
 *   - A plan is not re-entrant: concurrent calls on one plan from different
 *     streams race on its scratch.  phi handles are read-only after prepare.
 */
#ifndef CMBL_H
#define CMBL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cmbl_plan cmbl_plan;
typedef struct cmbl_phi cmbl_phi;
typedef int cmbl_status;

enum { CMBL_OK = 0, CMBL_EINVAL = 1, CMBL_EHIP = 2, CMBL_ENOMEM = 3, CMBL_ECAUSTIC = 4 };
enum { CMBL_F32 = 0, CMBL_F64 = 1 };
enum { CMBL_PHI_MAP = 0, CMBL_PHI_FOURIER = 1 };
enum { CMBL_FORWARD_LENSE = 1, CMBL_INVERSE_LENSE = -1 }; /* constants.py:7-8 */

/* Library version string. */
const char* cmbl_version(void);
/* Message of the last failing call on this thread. */
const char* cmbl_last_error(void);

/* Plan: geometry (N, pixel size in arcmin), arithmetic type and the largest
 * batch*ncomp it will be asked to process; allocates the stage scratch.
 * Replaces the static field metadata of Field (fields.py:15-22) plus the
 * per-call setup of rk4_solve (util.py:184-199).  N in {16,...,512}, 2^m. */
cmbl_status cmbl_plan_create(cmbl_plan** out, int N, double theta_pix_arcmin, int dtype,
                             int max_maps);
cmbl_status cmbl_plan_destroy(cmbl_plan* plan);

/* phi preparation: the five derivatives of phi (util.py:110-130, in float64)
 * and the table p(t_j) = (I + t_j Hess phi)^-1 grad phi for t_j = j/(2n),
 * j = 0..2n (lense_flow.py:372-381, 397-415), stored in the plan dtype.
 * phi is float64: MAP [N][N] or FOURIER [N][N/2+1] complex (lense_flow.py:299).
 * min_det (host out, may be NULL) receives min over pixels and t in [0,1] of
 * det(I + t Hess phi); with it the call synchronises `stream`, without it the
 * call only enqueues and the value is available later (cmbl_phi_min_det).
 * The temporaries live in the plan (allocated by its first prepare), so
 * preparations on one plan must not be issued concurrently from several host
 * threads.  Handle memory comes from a pool of released handles' memory when
 * a block of the same size is free (see cmbl_phi_release_async). */
cmbl_status cmbl_phi_prepare(cmbl_plan* plan, const void* phi, int phi_basis, int nsteps,
                             cmbl_phi** out, double* min_det, void* stream);
/* Batched phi preparation: `count` phis (contiguous [count][N][N] float64
 * MAP or [count][N][N/2+1] complex float64 FOURIER) prepared in one call:
 * their tables share one allocation, each preparation stage is one launch
 * over all of them, and the host waits once if min_det (an array of `count`)
 * is given, not at all if it is NULL.  out receives `count` handles, each
 * released on its own (the shared memory goes with the last).
 * Replaces the per-element preparation of a field batch with a phi per
 * element (lense_flow.py:297-302 applied to each; the reference recomputes
 * the derivatives on every call). */
cmbl_status cmbl_phi_prepare_batch(cmbl_plan* plan, const void* phis, int phi_basis, int nsteps,
                                   int count, cmbl_phi** out, double* min_det, void* stream);
/* Release a handle after waiting for the device (safe whatever streams used
 * it).  Its memory returns to the library's pool of handle memory. */
cmbl_status cmbl_phi_release(cmbl_phi* phi);
/* Release a handle without waiting: the caller has enqueued every use of the
 * handle on `stream` (or on streams already ordered before its current
 * point); the memory is reused by a later preparation only after an event
 * recorded here has completed.  A sampler whose phi changes every step thus
 * neither frees nor allocates device memory.  No reference counterpart: the
 * reference's arrays are garbage collected. */
cmbl_status cmbl_phi_release_async(cmbl_phi* phi, void* stream);
/* min over pixels and t in [0,1] of det(I + t Hess phi) of a prepared phi
 * (the caustic guard); waits for the device the first time it is asked. */
cmbl_status cmbl_phi_min_det(cmbl_phi* phi, double* out);
/* Device pointer of the phi derivative maps, float64 [5][N][N]
 * (phi_x, phi_y, phi_xx, phi_xy, phi_yy). */
const void* cmbl_phi_derivatives(const cmbl_phi* phi);

/* LenseFlow of `batch` fields of `ncomp` components each (1: T, 2: Q,U).
 * phis[b] (host array of handles) is the phi of batch element b; all must be
 * prepared with this plan and nsteps.  direction = +1 / -1, adjoint = 0 / 1
 * follow the truth table of lense_flow.py:317-322.  in may equal out.
 * Replaces lense_flow / primal_lense_flow (lense_flow.py:289-333). */
cmbl_status cmbl_lenseflow(cmbl_plan* plan, const cmbl_phi* const* phis, const void* in,
                           void* out, int batch, int ncomp, int nsteps, int direction,
                           int adjoint, void* stream);

/* LenseFlow that allocates nothing: the same computation, kernels and
 * results (bit for bit) as cmbl_lenseflow, but the stage spectra, the RK4
 * accumulator and the per-call phi table pointers live in a static device
 * array of CMBL_NOALLOC_SCRATCH_BYTES per GPU instead of the plan's scratch,
 * and the call makes no allocation, no host-to-device staging copy and no
 * timing events.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under stream capture, tight serving
 * loops).  The plan supplies geometry and twiddles only (it is created
 * beforehand, which also resolves the array's address on that device); the
 * caller passes the inputs and outputs.  A call needs
 * 3 * batch * ncomp * N^2 * sizeof(dtype) bytes (each region rounded to 256
 * bytes) plus 8 bytes per phi pointer; larger calls fail with CMBL_EINVAL
 * before any work.  Every call on one GPU uses that GPU's one array: calls
 * that overlap on a GPU (concurrent streams or host threads) race, so order
 * them on one stream.  The array keeps the last call's contents; nothing in
 * it must start at zero. */
#define CMBL_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
cmbl_status cmbl_lenseflow_noalloc(cmbl_plan* plan, const cmbl_phi* const* phis, const void* in,
                                   void* out, int batch, int ncomp, int nsteps, int direction,
                                   int adjoint, void* stream);

/* Backward rule of lense_flow_wrapper (lense_flow.py:10-51, 53-126, 156-252):
 * the joint RK4 of (f, delta f, delta phi) integrated from the transformed
 * field `lensed` (= lense_flow(f, phi, nsteps, direction)) and the cotangent
 * `cotangent` (both MAP [batch][ncomp][N][N]) with -direction.  Writes the raw
 * delta f (MAP, = L^T cotangent) into d_field and the delta-phi rate
 * accumulated in Fourier, complex [batch][N][N/2+1], into d_phi.  Both are
 * plain gradients: <d_field, w> and <irfft2(d_phi), v> are the directional
 * derivatives of <cotangent, L f> (undo_inner_product, fields.py:146-154, is
 * left to the caller).  d_field may equal cotangent. */
cmbl_status cmbl_lenseflow_vjp(cmbl_plan* plan, const cmbl_phi* const* phis, const void* lensed,
                               const void* cotangent, void* d_field, void* d_phi, int batch,
                               int ncomp, int nsteps, int direction, void* stream);

/* rfft2 / irfft2 of nmaps maps in the plan dtype with numpy semantics
 * (fields.py:158-172: fourier / map). */
cmbl_status cmbl_rfft2(cmbl_plan* plan, const void* x, void* X, int nmaps, void* stream);
cmbl_status cmbl_irfft2(cmbl_plan* plan, const void* X, void* x, int nmaps, void* stream);

/* Gaussian random fields in the FOURIER basis (simulate.py:100-113:
 * field_from_covar followed by rfft2) with the data model of load_sim fused
 * (simulate.py:699-702: mask * beam * lensed + rfft2(noise)).  For trial t and
 * component c (map m = t * ncomp + c, [ntrials][ncomp][N][N/2+1] in the plan
 * dtype):
 *   out[m] = (add ? mul[c] * add[m] : 0) + rfft2(irfft2(F_m, norm="ortho")),
 *   F_m = sqrt(cov[c] / 2) (a + i b),
 * with a, b standard normal per mode from Philox4x64-10 (key {seed, 0},
 * counter {kx (N/2+1) + ky, stream0 + (c << 32) + t, 0, 0}) and Box-Muller on
 * its first two outputs.  cov and mul are float64 [ncomp][N][N/2+1] device
 * arrays; mul may be NULL (1) when add is given and must be NULL without it.
 * The reference draws with jax.random; the streams differ, the distribution
 * and the per-mode construction are the same.  out may equal add. */
cmbl_status cmbl_grf_fourier(cmbl_plan* plan, const double* cov, const double* mul,
                             const void* add, void* out, int ntrials, int ncomp,
                             unsigned long long seed, unsigned long long stream0, void* stream);

/* ---- lensed Wiener filter -------------------------------------------------
 * Diagonal operators are real arrays [ncomp][N][N/2+1] in the plan dtype,
 * shared by all batch elements (DiagonalScalar / DiagonalEB,
 * matrix_operators.py:45-90).  Fields are EB (or T) FOURIER arrays.  With
 * ncomp = 3 (T, E, B: FlatS02, fields.py:107-116) the operators are BlockTEB
 * (matrix_operators.py:92-148): real [5][N][N/2+1] planes TT, TE, ET, EE, BB,
 * in the Fourier-mask mode only. */
typedef struct {
  const void* mb;       /* M B, per mode (Fourier-mask mode) */
  const void* bmcn;     /* B M Cn^-1, per mode (Fourier-mask mode) */
  const void* cfinv;    /* Cf^-1 per mode; NULL drops the prior term */
  const void* beam;     /* B per mode (pixel-mask mode) */
  const void* cninv;    /* Cn^-1 per mode (pixel-mask mode) */
  const void* mask_pix; /* real [N][N] pixel mask; NULL selects the Fourier-mask mode */
} cmbl_wf_ops;

/* gradf_logpdf (gradients.py:12-26):
 *   out = L^T B M^T Cn^-1 (d - M B L f) - Cf^-1 f       (EB FOURIER in and out)
 * with L = LenseFlow(phi, nsteps) forward and L^T its adjoint (the T maps and
 * the Q,U pairs of a T+EB field are lensed as two groups), and the EB<->QU /
 * map / fourier conversions of the reference.  f == NULL means f = 0 (the
 * forward LenseFlow of zeros is skipped); d == NULL means d = 0 (the CG
 * operator A, wiener_filter.py:45-49).  d is EB FOURIER in Fourier-mask mode,
 * a [batch][ncomp][N][N] MAP (T or Q,U) in pixel-mask mode.  out may not alias
 * f or d. */
cmbl_status cmbl_gradf_logpdf(cmbl_plan* plan, const cmbl_phi* const* phis, const void* f,
                              const void* d, const cmbl_wf_ops* ops, void* out, int batch,
                              int ncomp, int nsteps, void* stream);

/* The first half of gradf_logpdf for ncomp = 1 (T) or 2 (EB): the lensed maps
 * L f ([batch][ncomp][N][N] MAP, T or Q,U) and the cotangent of logpdf's data
 * term, map(eb2qu(B M^T Cn^-1 (d - M B L f))) (statistics.py:159; the cotangent
 * the reference's custom VJP receives, lense_flow.py:23-46).  With
 * cmbl_lenseflow_vjp from the lensed maps it gives grad_phi_logpdf
 * (gradients.py:29-37).  f is required; d == NULL means d = 0. */
cmbl_status cmbl_lensed_cotangent(cmbl_plan* plan, const cmbl_phi* const* phis, const void* f,
                                  const void* d, const cmbl_wf_ops* ops, void* lensed, void* cot,
                                  int batch, int ncomp, int nsteps, void* stream);

/* The last line of gradf_logpdf (gradients.py:25) from the adjoint-lensed
 * cotangent instead of from f: out = qu2eb(fourier(adj)) - Cf^-1 f, with
 * adj = L^T map(eb2qu(g)) [batch][ncomp][N][N] MAP (T or Q,U) -- the d_field
 * of cmbl_lenseflow_vjp run on cmbl_lensed_cotangent's two outputs -- f the
 * EB (or T) FOURIER field and cfinv Cf^-1 per mode [ncomp][N][N/2+1] in the
 * plan dtype (either NULL drops the prior term; the prior is applied per mode
 * to the Hermitian part of f, as cmbl_gradf_logpdf does).  So one forward
 * LenseFlow and one gradient ODE give both gradients of logpdf, as one
 * reverse pass of the reference's jax.grad does (gradients.py:29-37).
 * ncomp 1 (T) or 2 (EB); out may not alias f. */
cmbl_status cmbl_gradf_final(cmbl_plan* plan, const void* adj, const void* f, const void* cfinv,
                             void* out, int batch, int ncomp, void* stream);

/* FP64 inner products for nrhs right-hand sides (fields.py:136-142): out[k] =
 * Re sum conj(a_k) b_k w / N^2 over [ncomp][N][N/2+1].  a is complex double;
 * b is complex double (b_f64 = 1) or complex in the plan dtype.  out is a
 * device array of nrhs doubles. */
cmbl_status cmbl_dot_f64(cmbl_plan* plan, const void* a, const void* b, int b_f64, int nrhs,
                         int ncomp, double* out, void* stream);
/* One PCG update (wiener_filter.py:84-92) for nrhs systems, FP64 vectors:
 *   x += alpha p ; r -= alpha Ap ; z = pinv .* r ; rz = dot(r, z)
 * Ap is complex in the plan dtype (the operator's output); pinv is a double
 * [ncomp][N][N/2+1] diagonal, or for ncomp = 3 a BlockTEB [5][N][N/2+1] (TT,
 * TE, ET, EE, BB); alpha and rz are device arrays of nrhs doubles. */
cmbl_status cmbl_cg_update_xrz(cmbl_plan* plan, void* x, void* r, void* z, const void* p,
                               const void* Ap, const double* pinv, const double* alpha, int nrhs,
                               int ncomp, double* rz, void* stream);
/* p = z + beta p (wiener_filter.py:94-95); also writes p rounded to the plan
 * dtype into p_op when p_op is not NULL. */
cmbl_status cmbl_cg_update_p(cmbl_plan* plan, void* p, const void* z, const double* beta,
                             void* p_op, int nrhs, int ncomp, void* stream);

/* One PCG iteration after the operator (wiener_filter.py:84-95) with every
 * scalar kept on the device, for nrhs systems (FP64 vectors as above):
 *   pAp = dot(p, Ap) ; alpha = res / (pAp != 0 ? pAp : 1)
 *   x += alpha p ; r -= alpha Ap ; z = pinv .* r ; rz = dot(r, z)
 *   beta = rz / (res != 0 ? res : 1) ; p = z + beta p (p_op as in cmbl_cg_update_p)
 *   res = rz                                  (res: device array, updated in place)
 * for the right-hand sides with act[k] != 0 (act: device bytes, NULL = all);
 * the others get alpha = beta = 0 and keep res.  With rel (and bb, device
 * arrays) also rel = sqrt(dot(r, r) / bb).  The same float64 operations as
 * cmbl_dot_f64 + cmbl_cg_update_xrz + cmbl_cg_update_p with the scalars
 * formed on the host, in one call. */
cmbl_status cmbl_cg_step(cmbl_plan* plan, void* x, void* r, void* z, void* p, const void* Ap,
                         const double* pinv, void* p_op, double* res, const unsigned char* act,
                         const double* bb, double* rel, int nrhs, int ncomp, void* stream);

/* ---- serving executor ------------------------------------------------------
 * No counterpart in the reference (it serves one call at a time): a batched
 * LenseFlow of fixed capacity (max_batch requests of ncomp components, nsteps,
 * direction, adjoint) whose whole launch sequence is captured once into a
 * hipGraph (kernel nodes only) over the plan's staging buffer.  cmbl_exec_run
 * uploads the request pointers (device pointers to [ncomp][N][N] maps and
 * phi handles prepared for nsteps) and replays the graph on `stream`: gather
 * of the `count` inputs into staging, the RK4 stages, scatter of the results
 * to out_ptrs; slots past `count` lens zero maps and are not written back.
 * Executors use their plan's scratch: runs of executors that share a plan
 * must be ordered on one stream, and every executor is destroyed before its
 * plan.  Inputs and outputs must stay allocated until the run completes. */
typedef struct cmbl_exec cmbl_exec;
cmbl_status cmbl_exec_create(cmbl_exec** out, cmbl_plan* plan, int max_batch, int ncomp,
                             int nsteps, int direction, int adjoint);
cmbl_status cmbl_exec_run(cmbl_exec* ex, const cmbl_phi* const* phis, const void* const* in_ptrs,
                          void* const* out_ptrs, int count, void* stream);
cmbl_status cmbl_exec_destroy(cmbl_exec* ex);

/* Timing mode (no counterpart in the reference; measurement aid): when on,
 * every LenseFlow stage launch is bracketed by HIP events on the call's stream
 * and cmbl_lenseflow waits for its own work before returning.
 * cmbl_plan_get_timing writes out[6] = {column-pass ms, launches, row-pass ms,
 * launches, other ms, launches}, accumulated since the last reset. */
cmbl_status cmbl_plan_set_timing(cmbl_plan* plan, int enable);
cmbl_status cmbl_plan_get_timing(cmbl_plan* plan, double* out, int reset);

#ifdef __cplusplus
}
#endif
#endif /* CMBL_H */
