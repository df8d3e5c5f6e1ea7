/*
 * fgpe_c.h -- C ABI of the MI355X-native fGPE time-stepping library
 * (libfgpe.so, hand-written HIP kernels for gfx950).
 *
 * The reference has no C ABI: its hot path is C++ member functions of the
 * class chain Cuda <- BaseHamiltonian <- OneComponentGPSolver /
 * TwoComponentGPSolver <- plugin.  Each entry point below replaces one of
 * those members (cited per function); the reference-shaped C++ classes
 * (fgpe_amd/host) and the Python bindings call only this header.
 *
 * Conventions
 *   - Return value: 0 on success, a negative FGPE_E* code on failure; the
 *     message of the last failure on the calling thread is fgpe_last_error().
 *   - Complex arrays are interleaved (re, im) doubles in the reference's
 *     layout ID(x,y,z) = z*NX*NY + y*NX + x (hpp/header.hpp:20).
 *   - Host pointers are named host_*; dev_* arguments are device pointers
 *     (HIP allocations on the context's device), used on the context's
 *     stream.
 *   - A context owns its device memory and one stream; it is not
 *     thread-safe.  The caller owns the context handle.
 */
#ifndef FGPE_C_H
#define FGPE_C_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FGPE_OK 0
#define FGPE_EINVAL -1   /* bad argument                                   */
#define FGPE_EUNSUP -2   /* unsupported grid size / feature                */
#define FGPE_EHIP -3     /* HIP runtime error                              */
#define FGPE_ESTATE -4   /* call out of order (e.g. no k-space set)        */

/* precision of the bulk (FFT passes, right-hand side, stage states) */
#define FGPE_PREC_F64 0     /* everything FP64 (the plain drop-in form)       */
#define FGPE_PREC_MIXED 1   /* FP32 complex bulk; FP64 state, twiddles and
                               reductions                                    */

/* imaginary-time schemes */
#define FGPE_IMAG_RK4_RENORM 0  /* RK4 with c = -dt, renormalised each step  */
#define FGPE_IMAG_HEAVY_BALL 1  /* the reference's scheme,
                                   src/oneComponentGPSolver.cu:26-104         */

typedef struct fgpe_ctx fgpe_ctx;

/* Hamiltonian of one or two components.
     H psi_c = -1/2 lap psi_c + (V_c + sum_c' 4 pi a_cc' n_c' + Phi) psi_c
   Phi = F^-1[ Vt(k) F(n_total) ] / N with
     Vt = 12 pi a_dd ((k.d/|k|)^2 - 1/3)        (src/parallel.cu:326-352)
        + soft-core Vt of dimension sc_dim    (src/parallel.cu:371-477)
   The energy functional uses 2 pi a and Phi/2. */
typedef struct fgpe_hparams {
    int ncomp;          /* 1 or 2                                              */
    double a[3];        /* 1 comp: a[0]; 2 comp: a11, a22, a12 (code units)    */
    int quench_on;      /* 1: a(t) of d_quench_as (BEC_oneComponent.cu:167)   */
    double quench[6];   /* a0, a1, a2, t1, t2, t3                              */
    double a_dd;        /* dipolar length; 0 disables                          */
    double dip[3];      /* dipole direction                                    */
    double sc_U;        /* soft-core strength; 0 disables                      */
    double sc_R;        /* soft-core radius                                    */
    int sc_dim;         /* 1, 2 or 3: which soft-core transform                */
    double vpx[2];      /* moving-frame term -vpx_c p_x / (2 pi) per component,
                           i.e. -vpx_c F^-1[k_x F(psi_c)] / N (alg_addVPx,
                           src/baseHamiltonian.cu:198-211; trotate:
                           vpx = 2 pi omega / (omho aho), BEC_oneComponent.cu:98-101,
                           BEC_twoComponent.cu:125-129); 0 disables.  It enters
                           where the reference's alg_calcHpsi does: heavy ball,
                           observables and H applications, and only stage 4 of
                           the one-component RK4 step (stages 1-3 use
                           alg_calcHpsiMU, src/oneComponentGPSolver.cu:170-187;
                           the two-component RK4 uses only alg_calcHpsiMU). */
    double user[8];     /* parameters of the user pointwise term, read by its
                           functor (fgpe_amd/csrc/user_term.hpp)             */
    int user_id;        /* user pointwise term of this context: 0 none, else
                           one of the ids fgpe_user_term_ids lists         */
} fgpe_hparams;

/* Custom Hamiltonians (the plugin kernels SingleTemplate_template /
   SingleTemplate_MU_template, hamiltonians/single_template.cu:333-373, that
   a user edits; main.cu:41-69 picks one of several plugins by name): the
   library compiles a list of user functors, each with an id, into the fused
   pointwise stage of every H application; fgpe_hparams::user_id selects one
   per context.  It adds w psi_c to H psi_c and w_en psi_c to the energy
   operator, w from (t, grid index, component, psi_c, densities, user[8]).
   Editing fgpe_amd/csrc/user_term.hpp and rebuilding the library changes
   the list; the shipped library has none.  *enabled = 1 when this build has
   at least one; fgpe_user_term_ids writes up to cap ids and *count, the
   number compiled.  Adjoint gradients and batched groups reject a context
   with a user term (its derivative is not known to the library; group
   launches run at a fixed time). */
int fgpe_user_term(int* enabled);
int fgpe_user_term_ids(int* ids, int cap, int* count);

/* ---- context (replaces Cuda::Cuda, src/parallel.cu:14-113) ------------- */
/* Grid extents: nx and, when above 1, ny and nz are each one of 16, 32, 48,
   64, 96, 128, 192, 256, 384, 512, 1024 (other lengths return FGPE_EUNSUP
   with the list); a 2D grid has nz = 1 and a 1D grid ny = nz = 1
   (src/parallel.cu:67-110, src/params.cu:145-154). */
int fgpe_ctx_create(int nx, int ny, int nz, int precision, fgpe_ctx** out);
int fgpe_ctx_destroy(fgpe_ctx* ctx);
const char* fgpe_last_error(void);
/* Run on the given HIP stream from now on (NULL selects the HIP null
   stream).  A context starts on a non-blocking stream of its own. */
int fgpe_ctx_set_stream(fgpe_ctx* ctx, void* hip_stream);
int fgpe_ctx_sync(fgpe_ctx* ctx);
/* Bytes of device memory the context holds. */
int fgpe_ctx_device_bytes(fgpe_ctx* ctx, size_t* bytes);

/* k axes in cycles per length and the volume element
   (BaseHamiltonian::InitializeKspace, src/baseHamiltonian.cu:53-82; DV from
   src/params.cu:213). */
int fgpe_set_kspace(fgpe_ctx* ctx, const double* host_kx, const double* host_ky,
                    const double* host_kz, double dv);
/* Hamiltonian parameters (the plugin constructor's d_h_params,
   hamiltonians/BEC_oneComponent.cu:10-15, plus the dipolar/soft-core
   kernels). */
int fgpe_set_hamiltonian(fgpe_ctx* ctx, const fgpe_hparams* hp);

/* State transfer (the H2D/D2H copies of runRealTimeEvol/runImagTimeEvol,
   src/oneComponentGPSolver.cu:32-34,127-129,211). comp = 0 or 1. */
int fgpe_upload_psi(fgpe_ctx* ctx, int comp, const double* host_psi);
int fgpe_download_psi(fgpe_ctx* ctx, int comp, double* host_psi);
int fgpe_upload_vext(fgpe_ctx* ctx, int comp, const double* host_vext);
/* The same from and to device memory, ordered on the context's stream (no
   host round trip; for callers whose arrays already live on the device,
   such as the autograd function of the adjoint).  The trap form takes a real
   FP64 potential (N doubles). */
int fgpe_upload_psi_dev(fgpe_ctx* ctx, int comp, const void* dev_psi);
int fgpe_download_psi_dev(fgpe_ctx* ctx, int comp, void* dev_psi);
int fgpe_upload_vext_real_dev(fgpe_ctx* ctx, int comp, const double* dev_vext);
/* Device address of component comp's FP64 state (N complex doubles). */
int fgpe_psi_device_ptr(fgpe_ctx* ctx, int comp, void** dev_psi);

/* ---- initial state and trap built on the device ------------------------ */
/* Cell centres x_i = -xmax + i dx + dx/2 per axis (Psi::getX/getY/getZ,
   src/psi.cu:134-150; Vext::getX/getY/getZ, src/vext.cu:182-199). */
typedef struct fgpe_geom {
    double xmax[3];
    double dx[3];
} fgpe_geom;

enum { FGPE_INIT_HARMONIC = 0, FGPE_INIT_TUBE = 1, FGPE_INIT_VOID = 2 };
enum {
    FGPE_LATTICE_1D = 0,          /* Psi::force1DLattice        src/psi.cu:179-193 */
    FGPE_LATTICE_2D_TRIANGULAR,   /* Psi::force2DTriangularLattice :196-213 */
    FGPE_LATTICE_2D_SQUARE,       /* Psi::force2DSquareLattice  :217-232 */
    FGPE_LATTICE_3D_HEX,          /* Psi::force3DHEXLattice     :236-262 */
    FGPE_LATTICE_3D_FCC,          /* Psi::force3DFCCLattice     :265-297 */
    FGPE_LATTICE_3D_BCC,          /* Psi::force3DBCCLattice     :301-328 */
    FGPE_LATTICE_3D_SC            /* Psi::force3DSCLattice      :331-347 */
};

/* psi_comp = (1 + noise) * envelope at every point: envelope
   exp(-0.05 (x^2 + y^2 + z^2)) on x/ax (HARMONIC, Psi::initHarmonic,
   src/psi.cu:100-116), the same without x (TUBE, Psi::initTube :80-97) or
   1 (VOID, Psi::initVoid :119-131); noise = (rand_re u1, rand_im u2) with u
   uniform on [-1, 1).  The reference draws u from a sequential
   std::mt19937 stream; here u comes from the counter-based Philox4x32-10
   generator (Salmon et al., SC'11) keyed by seed, counter (point index,
   stream 0), so a point's noise does not depend on the others.  With zero
   amplitudes the result is the reference's.  Not normalised. */
int fgpe_psi_init(fgpe_ctx* ctx, int comp, int kind, const fgpe_geom* g, const double ax[3],
                  double rand_re, double rand_im, unsigned long long seed);
/* psi *= 1 + (amp_re u1, amp_im u2), u from Philox stream `stream` (>= 1;
   Psi::addNoise, src/psi.cu:167-176). */
int fgpe_psi_add_noise(fgpe_ctx* ctx, int comp, double amp_re, double amp_im, unsigned long long seed,
                       unsigned int stream);
/* psi *= exp(i 2 pi ncirc x / (2 xmax_x)) (Psi::imprintVortexTube,
   src/psi.cu:155-164). */
int fgpe_psi_imprint_vortex_tube(fgpe_ctx* ctx, int comp, const fgpe_geom* g, double ncirc);
/* psi += force(x - off) * sqrt(npart / (dx dy dz)) with the lattice's
   force (FGPE_LATTICE_*); the reference then normalises with component
   1's particle number, which the caller does with fgpe_psi_normalize. */
int fgpe_psi_force_lattice(fgpe_ctx* ctx, int comp, int lattice, const fgpe_geom* g, double amp, double k,
                           const double off[3], double npart);
/* psi *= sqrt(npart / (sum |psi|^2 dx dy dz)); host_norm (may be NULL)
   receives sum |psi|^2 dx dy dz afterwards (Psi::normalize,
   src/psi.cu:351-372). */
int fgpe_psi_normalize(fgpe_ctx* ctx, int comp, double npart, const fgpe_geom* g, double* host_norm);

/* Trap: the real FP64 potential is assembled on the device and enters the
   Hamiltonian at fgpe_vext_commit.  HARMONIC 0.5 ((x/ax)^2/ax^2 + ...)
   (Vext::initHarmonic, src/vext.cu:72-85), TUBE without x (Vext::initTube
   :60-70), VOID 0 (:112-118). */
int fgpe_vext_init(fgpe_ctx* ctx, int comp, int kind, const fgpe_geom* g, const double ax[3]);
/* Box trap 0.5 V prod_a (tanh(B (x_a - w_a)) - tanh(B (x_a + w_a))) on the
   raw cell centres (Vext::initBox, src/vext.cu:88-109; w_a = xbox/aho ...). */
int fgpe_vext_init_box(fgpe_ctx* ctx, int comp, const fgpe_geom* g, double V, double B, const double w[3]);
/* V += A cos(2 pi n x / (2 xmax_x)) (Vext::addProtocolPotential,
   addWeightedProtocolPotential with its host-side amplitude, and
   addOpticalLattice, src/vext.cu:122-179). */
int fgpe_vext_add_cos(fgpe_ctx* ctx, int comp, const fgpe_geom* g, double A, double n);
/* Make the assembled potential component comp's trap (rounded to the
   bulk precision); host_vext (may be NULL) receives it as N complex
   doubles, for the Vext output file. */
int fgpe_vext_commit(fgpe_ctx* ctx, int comp, double* host_vext);

/* ---- building blocks on device arrays ---------------------------------- */
/* 3D C2C transform of N complex values, unnormalised; direction -1 forward
   (CUFFT_FORWARD), +1 inverse (cufftExecZ2Z, src/baseHamiltonian.cu:89,95).
   in == out is allowed.  _f32 works on interleaved floats. */
int fgpe_fft3d_c2c_f64(fgpe_ctx* ctx, const void* dev_in, void* dev_out, int direction);
int fgpe_fft3d_c2c_f32(fgpe_ctx* ctx, const void* dev_in, void* dev_out, int direction);
/* out = F^-1[(19.739208802/N) |k|^2 F psi] = -lap psi / 2
   (alg_Laplace + Laplace, src/baseHamiltonian.cu:86-96, src/parallel.cu:194-205) */
int fgpe_laplacian_f64(fgpe_ctx* ctx, const void* dev_psi, void* dev_out);

/* H applied to the current state at time t (alg_calcHpsi,
   hamiltonians/BEC_oneComponent.cu:92-103): dev_hmu, dev_hen receive
   ncomp*N complex doubles each (dev_hen may be NULL). */
int fgpe_apply_h(fgpe_ctx* ctx, double t, void* dev_hmu, void* dev_hen);
/* norm = sum |psi_c|^2 DV, mu = sum Re(psi* H psi) DV / norm,
   E = sum Re(psi* Hen psi) DV / norm, for each component c:
   out[3*c + {0,1,2}] (alg_calcNorm + alg_calc2Observables,
   src/baseHamiltonian.cu:173-179,230-236). */
int fgpe_observables(fgpe_ctx* ctx, double t, double* host_out);

/* Diagnostics on component comp of the current state:
   px  = DV sum_k kx |F psi|^2, which equals the reference's
         DV sum Re(psi* F^-1[kx F psi]) (alg_calcPx, src/baseHamiltonian.cu
         :184-195, with BECVPx and CalculateObservable)
   cos = DV sum cos(2 pi x / L) |psi|^2 with host_x the NX cell centres
         (CalcAverageCos, src/parallel.cu:535-561, through alg_calcCos). */
int fgpe_calc_px(fgpe_ctx* ctx, int comp, double* host_px);
int fgpe_calc_cos(fgpe_ctx* ctx, int comp, const double* host_x, double L, double* host_out);

/* ---- time stepping ----------------------------------------------------- */
/* nsteps real-time RK4 steps from time t0 (runRealTimeEvol's loop body,
   src/oneComponentGPSolver.cu:150-208; two components:
   src/twoComponentGPSolver.cu:142-203). */
int fgpe_rk4_steps(fgpe_ctx* ctx, double dt, double t0, int nsteps);
/* nsteps imaginary-time steps (scheme FGPE_IMAG_*).  host_npart: target
   norm of each component (ncomp doubles).  eps > 0 stops heavy-ball at |E - E_prev| / E_prev < eps
   (src/oneComponentGPSolver.cu:84-88; two components: |dE_c| < eps for
   both, src/twoComponentGPSolver.cu:94).  host_hist (3*ncomp*nsteps doubles
   or NULL) gets norm, mu, E per component per step (heavy ball); the RK4
   scheme evaluates no energy operator and fills only the norm, with NaN in
   the mu and E slots; *steps_done the number run.  A heavy-ball call
   continues the previous one (momentum, mu, E_prev, convergence) unless the
   state or the Hamiltonian was set in between.  Time continues across calls
   as well: iteration j of a call runs at t = (m + j) dt, m the iterations
   run since the state or Hamiltonian was set or a heavy-ball run started
   afresh (PARS[T] = iteration * dt, src/oneComponentGPSolver.cu:50), so a
   run split into several calls matches one call. */
int fgpe_imag_steps(fgpe_ctx* ctx, int scheme, double dt, double beta, const double* host_npart,
                    double eps, int nsteps, double* host_hist, int* steps_done);
/* fgpe_imag_steps for callers that cannot allocate device memory (code under
   HIP graph capture, serving loops, harnesses that forbid hipMalloc).  No
   counterpart in the reference.  Same arguments, results and errors, but the
   per-iteration history goes to a fixed device array of the library instead
   of a buffer allocated on first use: the call makes no hipMalloc / hipFree.
   The array holds the history of at most FGPE_NOALLOC_MAX_STEPS iterations
   (3 * 2 * FGPE_NOALLOC_MAX_STEPS doubles, 768 KiB on every device the
   library runs on, kept until the process ends); a call asking for a
   history of more iterations returns FGPE_EINVAL (without a history any
   nsteps is accepted).  Every context on one device shares that array: calls
   that overlap on one device (other host threads, other streams) race. */
#define FGPE_NOALLOC_MAX_STEPS 16384
int fgpe_imag_steps_noalloc(fgpe_ctx* ctx, int scheme, double dt, double beta, const double* host_npart,
                            double eps, int nsteps, double* host_hist, int* steps_done);
/* *converged = 1 once a heavy-ball run has met its convergence test. */
int fgpe_imag_converged(fgpe_ctx* ctx, int* converged);

/* ---- adjoint (no counterpart in the reference) --------------------------
   Gradients of a real objective J(psi(T)) of nsteps real-time RK4 steps with
   respect to the trap V, the contact lengths and the strength of the k-space
   kernel (a_dd or the soft-core U), by the discrete adjoint of the RK4 step:
   one or two components, real V, the dipolar kernel, the soft-core kernel
   or both (their direction, radius and dimension stay fixed).  With the
   one-component a(t) ramp on (fgpe_hparams::quench_on) the ramp is part of the
   differentiated dynamics and the derivative in a is along a uniform shift of
   a(t).
   fgpe_adjoint_forward runs the steps from the current state and keeps a
   checkpoint every ckpt_every steps (about sqrt(nsteps) keeps the memory at
   O(sqrt(nsteps)) states); psi(T) is then the current state.
   fgpe_adjoint_backward takes dev_lamT = dJ/dRe(psi) + i dJ/dIm(psi) at T
   (ncomp x N complex doubles, component-major, device), recomputes each
   segment from its checkpoint (keeping, per step of the segment, the FP64
   state and the three RK stage states in the context's precision: device
   memory of 4 ckpt_every (FP64) or 2.5 ckpt_every (mixed) FP64 states per
   component besides the checkpoints), and writes dJ/dpsi(0) in the same convention
   to dev_lam0 (or NULL), dJ/dV_c to dev_gradV (ncomp x N doubles) and
   host_grad (5 doubles): [0] dJ/da (one component) or dJ/da_11, [1] dJ/da_dd,
   [2] dJ/dU, [3] dJ/da_22, [4] dJ/da_12 (0 where a parameter is absent).
   The state is left at the start of the last segment. */
int fgpe_adjoint_forward(fgpe_ctx* ctx, double dt, double t0, int nsteps, int ckpt_every);
int fgpe_adjoint_backward(fgpe_ctx* ctx, const void* dev_lamT, void* dev_lam0, double* dev_gradV,
                          double* host_grad);

/* ---- instrumentation (no counterpart in the reference, which prints only
   an "Average iteration time", hamiltonians/BEC_oneComponent.cu:279) ---- */
#define FGPE_KCLASS_XSTAGE 0   /* fused x pass (pointwise H + RK stage)      */
#define FGPE_KCLASS_STRIDED 1  /* y (or z) forward/inverse pass               */
#define FGPE_KCLASS_LAST 2     /* last axis forward + k multiply + inverse    */
#define FGPE_KCLASS_FFTX 3     /* standalone x pass                           */
#define FGPE_KCLASS_OTHER 4    /* reductions and small kernels                */
#define FGPE_NKCLASS 5
/* on != 0: bracket every launch with HIP events on the context's stream and
   accumulate per kernel class; on == 0 stops.  Enabling clears the record. */
int fgpe_timing_enable(fgpe_ctx* ctx, int on);
/* Synchronises, then writes out[2*k] = total ms, out[2*k+1] = launches for
   each class k < FGPE_NKCLASS, and clears the record. */
int fgpe_timing_read(fgpe_ctx* ctx, double* host_out);

/* ---- plan variants and the autotuner (no counterpart in the reference) ----
   Every kernel class has two compiled plans per pencil length: variant 0
   (16 elements per thread on long pencils, 128-thread x blocks, 128-byte row
   segments; FP64 y and last-axis passes of length 256: 256-byte rows and
   exchanges split into real and imaginary halves) and variant 1 (8 elements
   per thread where a second radix split exists, 256-thread x blocks,
   256-byte row segments).  fgpe_autotune times
   both for the x-stage, strided and last-axis classes on the context's own
   buffers (reps launches each; the state is preserved) and keeps the faster:
   the x stage alone, the y and last-axis passes in whole RK stages (2 * reps;
   a pass's speed depends on what the pass before it left in the memory-side
   cache).
   It then times whole RK stages: the 16 block orders (fgpe_set_order), on 3D
   grids for each choice of the k-space multiply axis (fgpe_set_kmul_axis,
   with that choice's pass variants tuned first and once more in the chosen
   order, the order again if a variant changed), keeping the fastest, and
   last the y / last / y sequence split into 1, 2 or 4 column slabs (a slab's
   three passes run back to back while it is in the memory-side cache).
   The x-stage class has a third plan, variant 2 (4 elements per thread in
   radix-4 stages, four pencils per block), on nx = 128 and 256; elsewhere
   variant 2 runs variant 1.
   host_ms (10 doubles or NULL) gets the times per class and variant 0/1 of
   the chosen configuration (6: strided, last-axis, x stage; the passes'
   times are those of whole RK stages with that variant), then per slab
   count (3; 0 where not applicable), then the x stage's variant 2 (0 where
   the length has none).  Needs the Hamiltonian, trap and state set.
   fgpe_set_variant takes 0 or 1, and 2 for FGPE_KCLASS_XSTAGE. */
int fgpe_autotune(fgpe_ctx* ctx, int reps, double* host_ms);
int fgpe_set_variant(fgpe_ctx* ctx, int kclass, int axis, int variant);
int fgpe_get_variant(fgpe_ctx* ctx, int kclass, int axis, int* variant);
/* Column slabs of the y / last / y sequence: 1, 2 or 4 (nx divisible). */
int fgpe_set_slabs(fgpe_ctx* ctx, int nslab);
int fgpe_get_slabs(fgpe_ctx* ctx, int* nslab);
/* Block order of the passes of a stage: bit 0 y forward, bit 1 last axis,
   bit 2 y inverse, bit 3 x stage; a set bit runs that pass's blocks in
   reverse so it starts on what the pass before it wrote last.  Results do
   not depend on it.  fgpe_autotune picks it last (16 combinations timed as
   whole RK stages). */
int fgpe_set_order(fgpe_ctx* ctx, int mask);
int fgpe_get_order(fgpe_ctx* ctx, int* mask);
/* The axis whose pass carries the k-space multiply (the forward FFT, the
   kinetic and k-space-kernel factors and the inverse FFT in one pass): on a
   3D grid 2 (z, the default; y runs the strided forward and inverse passes)
   or 1 (y; z runs them), on a 2D grid 1, on a 1D grid 0.  fgpe_autotune picks
   it on 3D grids (each choice's pass variants tuned, one RK stage timed per
   choice).  The two 3D choices agree to rounding (the y and z transforms run
   in the other order).  Batched groups always use z. */
int fgpe_set_kmul_axis(fgpe_ctx* ctx, int axis);
int fgpe_get_kmul_axis(fgpe_ctx* ctx, int* axis);

/* ---- batched sweeps (no counterpart in the reference, which runs one
   instance per process: src/oneComponentGPSolver.cu:26-104) ----------------
   A group is n contexts of one shape, precision, stream and Hamiltonian kind
   (one component, the same k-space kernel configuration; a, a_dd, the
   soft-core term, trap and state may differ).  Its kernels run as one launch
   per kernel class with a table of per-instance arguments, so small grids
   fill the GPU together.  The contexts stay usable on their own; destroy the
   group before any of its contexts. */
typedef struct fgpe_group fgpe_group;
int fgpe_group_create(fgpe_ctx* const* ctxs, int n, fgpe_group** out);
int fgpe_group_destroy(fgpe_group* group);
/* Times both plan variants of each kernel class as batched launches and sets
   the faster on every context (host_ms: 10 doubles or NULL, as fgpe_autotune;
   groups run unslabbed, the slab slots stay 0), then the block order
   (fgpe_set_order's bits; a batched pass run reversed also takes its
   instances in reverse) timed as whole batched iterations.  A heavy-ball run in progress
   ends here (its pending renormalisation is applied); the next
   fgpe_group_imag_steps starts a fresh run from the normalised states. */
int fgpe_group_autotune(fgpe_group* group, int reps, double* host_ms);
/* nsteps heavy-ball iterations of every instance (same scheme and stopping
   test as fgpe_imag_steps, per instance), enqueued without a host
   synchronisation; a converged instance does nothing in later iterations.
   Continues a previous group run unless a state or Hamiltonian was set. */
int fgpe_group_imag_steps(fgpe_group* group, double dt, double beta, const double* host_npart, double eps,
                          int nsteps);
/* Iterations each instance has run since the run started or was continued
   (iters, n ints or NULL) and whether it has converged (converged, n ints or
   NULL); synchronises, leaves the run state alone. */
int fgpe_group_imag_status(fgpe_group* group, int* iters, int* converged);
/* Ends a run: applies the pending renormalisation to every state, then
   reports as fgpe_group_imag_status. */
int fgpe_group_imag_finish(fgpe_group* group, int* iters, int* converged);

/* ---- captured launch sequences -------------------------------------------
   fgpe_graph_capture_imag prepares the groups (as fgpe_group_imag_steps does
   on its first call), then captures `iters` heavy-ball iterations of all of
   them as one hipGraph on their common stream (not the null stream), each
   group on a branch of its own (the groups are independent, so their
   kernels may run side by side).  Each
   fgpe_graph_launch replays those iterations; the host polls with
   fgpe_group_imag_status between launches and ends the run with
   fgpe_group_imag_finish.  The graph keeps its own copy of the argument
   tables; destroy it before its groups' contexts. */
typedef struct fgpe_graph fgpe_graph;
int fgpe_graph_capture_imag(fgpe_group* const* groups, int ngroups, double dt, double beta,
                            const double* host_npart, double eps, int iters, fgpe_graph** out);
int fgpe_graph_launch(fgpe_graph* graph);
int fgpe_graph_destroy(fgpe_graph* graph);

#ifdef __cplusplus
}
#endif
#endif
