/*
 * jaqsi_amd C-ABI: the drop-in boundary of the MI355X simulation backend.
 *
 * The reference's seam is the Python function
 *   simulate_and_measure(tape, n_qubits, type, obs, use_density, shots, key,
 *                        initial_state, adjoint)
 * (jaqsi/simulation.py:658-732), reached from Script.execute
 * (jaqsi/script.py:226-335, 420-434).  Everything below it is replaced here:
 *
 *   jq_plan_create        <- _compile / _fuse / _fuse_runs / _compile_mixed
 *                            (simulation.py:214-381): the recorded tape, lowered
 *                            to the jq_op/jq_member list below, is fused and
 *                            scheduled into tile passes.
 *   jq_sv_forward         <- _run + measure_state / measure_density
 *                            (simulation.py:400-473, 735-857)
 *   jq_sv_adjoint         <- _adjoint_expval (simulation.py:546-655), and for
 *                            density matrices the checkpointed adjoint that the
 *                            reference does not have (it tapes every rho).
 *   jq_pulse_propagate    <- Evolution magnus2 solve + unitary_cost_fn
 *                            (evolution.py:194-263, qoc.py:261-343)
 *   jq_last_error         <- the Python exceptions of the reference (ValueError /
 *                            TypeError / RuntimeError), mapped by status code.
 *
 * Conventions
 *  - Bit order: wire 0 is the most significant bit of the flat index
 *    (simulation.py:794-796).  A density matrix of n qubits is simulated as a
 *    vector of 2n "wires": ket wires 0..n-1, bra wires n..2n-1 (row-major rho).
 *  - All device pointers are caller-owned; the library keeps none of them
 *    after a call returns.  Persistent state lives only in jq_plan.
 *  - Every compute entry point takes the HIP stream to launch on and is
 *    asynchronous on it.
 *  - Cotangents follow JAX's convention (no conjugation), i.e. the value the
 *    reference's custom_vjp returns; PyTorch's convention is its conjugate.
 */
#ifndef JAQSI_AMD_H
#define JAQSI_AMD_H

#ifndef __HIPCC_RTC__
#include <stdint.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define JQ_OK 0
#define JQ_ERR_ARG 1         /* ValueError in the reference */
#define JQ_ERR_UNSUPPORTED 2 /* NotImplementedError / TypeError */
#define JQ_ERR_ALLOC 3       /* device allocation failure */
#define JQ_ERR_DEVICE 4      /* HIP runtime error */
#define JQ_ERR_NONFINITE 5   /* non-finite or singular result */

#define JQ_MAX_WIRES 6

/* member kinds: one primitive factor of a recorded gate */
#define JQ_M_ROT 0    /* controlled(cos(t/2) I - i sin(t/2) P_word, nctrl) */
#define JQ_M_PHASE 1  /* controlled(diag(1, e^{i t}), nctrl)                */
#define JQ_M_MATRIX 2 /* explicit matrix from the matrix pool               */

/* member flags */
#define JQ_MF_CONJ 1    /* use the complex conjugate of the factor          */
#define JQ_MF_SUPEROP 2 /* use kron(m, conj(m)) on (ket wires, bra wires)   */
#define JQ_MF_GRAD 4    /* an explicit matrix whose cotangent is requested  */
#define JQ_MF_UNITARY 8 /* an explicit matrix known to be unitary (a library
                           gate's own matrix): lets fold_tail drop it when diagonal */

/* op structure flags (GateStructure, operations.py:176-185) */
#define JQ_S_DIAGONAL 1
#define JQ_S_PERMUTATION 2
#define JQ_S_NONUNITARY 4 /* channel superoperator (not unitary)                */

/* A non-unitary block whose condition number (the product of its ops'
 * jq_op.cond) exceeds this bound is not un-applied by its inverse in the
 * adjoint sweep: it runs in a pass of its own, whose backward reads the
 * checkpointed input state instead.  Singular channels (cond = inf, e.g.
 * BitFlip(0.5), AmplitudeDamping(1)) always take that path. */
#define JQ_COND_MAX 4.0

typedef struct jq_member {
  int32_t kind;   /* JQ_M_* */
  int32_t flags;  /* JQ_MF_* */
  int32_t nctrl;  /* leading control wires of this factor */
  int32_t word;   /* Pauli word, 2 bits per target wire (I0 X1 Y2 Z3), first target highest */
  int32_t param;  /* index into the angle vector, or -1 for a constant */
  int32_t matrix; /* JQ_M_MATRIX: element offset of the matrix in the pool (complex128) */
  double value;   /* the angle when param == -1 */
} jq_member;

typedef struct jq_op {
  int32_t nwires;
  int32_t wires[JQ_MAX_WIRES];
  int32_t structure;    /* JQ_S_* */
  int32_t controls;     /* GateStructure.controls */
  int8_t perm[64];      /* GateStructure.permutation when JQ_S_PERMUTATION */
  int32_t member_begin; /* members of this op: [member_begin, member_end) */
  int32_t member_end;
  double cond;          /* JQ_S_NONUNITARY: 2-norm condition number of the op's
                           matrix (inf if singular or unknown); ignored otherwise */
} jq_op;

/* observable kinds for the fused reductions */
#define JQ_OBS_NONE 0
#define JQ_OBS_SV_DIAG 1 /* <psi|O|psi>, O = diag(d0,d1) on one wire (simulation.py:772-801) */
#define JQ_OBS_DM_DIAG 2 /* Tr(O rho), O = diag(d0,d1) on one ket wire  (simulation.py:850-852) */
/* <psi|O|psi>, O|x> = d[parity(x & mask)] |x>: a Pauli-Z word (d0 = 1, d1 = -1) on
 * the wires of `mask`.  For this kind obs_wires holds TWO int32 per observable,
 * the low and high 32 bits of the wire mask (bit w = wire w).  The GPU Script
 * path measures Pauli words / Hermitian observables through it after basis
 * rotations (simulation.py:806-809 is the dense fallback it replaces). */
#define JQ_OBS_SV_PARITY 3

typedef struct jq_plan_desc {
  int32_t n_wires;        /* qubits of the simulated vector (2n for a density matrix) */
  int32_t n_ops;
  const jq_op* ops;
  int32_t n_members;
  const jq_member* members;
  int32_t n_params;       /* length of the angle vector of one circuit */
  int32_t n_matrix_elems; /* complex elements of the explicit-matrix pool of one circuit */
  int32_t dtype;          /* 0 = complex64, 1 = complex128 */
  int32_t tile_bits;      /* 0 = choose; else tile size 2^tile_bits (<= n_wires); complex64 also takes 14
                           * (plan-specialised kernels only, no 4x4 or dense 3-wire blocks) */
  int32_t checkpoints;    /* 1 = keep every pass output for the backward sweep */
  /* Tuning knobs, 0 = the built-in choice (jaqsi_amd.autotune times the
   * candidates per circuit structure and batch and caches the winner):
   *   low_bits       lowest global bits every tile holds, 3..6 (contiguous
   *                  runs of 2^low_bits amplitudes per load; fewer leave more
   *                  tile bits to the gates and can cut passes)
   *   variant_flags  JQ_TUNE_* bits selecting kernel variants */
  int32_t low_bits;
  int32_t variant_flags;
  /* 1 = the plan is only measured through statevector observables
   * (JQ_OBS_SV_DIAG / JQ_OBS_SV_PARITY expectation values and their
   * gradients): trailing parameter-free blocks that permute basis states
   * affinely (X, CX, SWAP and their products) or are diagonal (CZ, S, T,
   * fixed phases), with no later op on their wires, are not simulated; the
   * observables are measured through them instead (a Pauli-Z word after the
   * tail is a signed Pauli-Z word before it).  Expectation values and
   * gradients are those of the full circuit; `state` outputs hold the vector
   * before the folded tail, and a final-vector cotangent (lam_seed) or
   * density-matrix observables are refused (JQ_ERR_UNSUPPORTED).
   * jq_plan_tables(6) reports the fold.  No reference counterpart: the
   * reference simulates every gate (simulation.py:400-409). */
  int32_t fold_tail;
} jq_plan_desc;

#define JQ_TUNE_NO_PERSIST_BWD 1 /* backward: one tile per workgroup, no persistent loop with prefetch */
#define JQ_TUNE_NO_DMA_FWD 2     /* complex64 forward: loads through registers, not direct to LDS */
#define JQ_TUNE_NARROW_BWD 4     /* complex64 plan-specialised backward: 8 amplitudes x 1024 threads */
#define JQ_TUNE_BWD16 8          /* complex64 plan-specialised backward: 16 amplitudes per thread (13-bit tiles:
                                    512 threads, persistent) instead of 32 (13-bit tiles: 256 threads, two
                                    workgroups per CU) */

typedef struct jq_plan jq_plan;

/* Lower, fuse and schedule one circuit structure; upload its program.
 * Replaces simulation._compile/_compile_mixed (simulation.py:338-381).
 * A plan keeps one set of workspaces (block matrices, checkpoints, partials)
 * per stream it has run on, so calls on different streams may overlap; the
 * sets live until jq_plan_destroy. */
int jq_plan_create(const jq_plan_desc* desc, jq_plan** out);
int jq_plan_destroy(jq_plan* plan);

/* Scheduling summary: passes, substeps, blocks, cotangent reals. */
int jq_plan_info(const jq_plan* plan, int32_t* n_passes, int32_t* n_substeps, int32_t* n_blocks,
                 int32_t* n_grad_reals, int32_t* tile_bits);

/* Host copy of the scheduled program (int32 words) for inspection. */
int jq_plan_program(const jq_plan* plan, int32_t* out, int64_t capacity, int64_t* size);

/* Host copies of the plan's tables: which = 0 block table, 1 member list,
 * 2 member table, 3 cotangent slot map, 4 backward program of the prebuilt
 * kernels, 5 backward program of the plan-specialised kernels, 6 the folded
 * tail (fold_tail): [ops folded, flip low/high 32 bits, then per wire w the
 * low/high 32 bits of the wire mask whose parity is wire w's final bit].  No
 * reference counterpart: these let tests execute the scheduled program on
 * the CPU. */
int jq_plan_tables(const jq_plan* plan, int32_t which, int32_t* out, int64_t capacity, int64_t* size);

/* Plan-specialised pass kernels.  The prebuilt kernels interpret the
 * program; for one plan the same gate sequence can also be compiled at run
 * time (hipRTC) into straight-line kernels, one per pass and direction,
 * which give the same results with less instruction overhead per gate.
 * No reference counterpart: the reference's jax.jit tracing per structure
 * (script.py:389-471) plays this role there.
 *   mode JQ_JIT_OFF       run the prebuilt kernels;
 *        JQ_JIT_ASYNC     compile on a worker thread, switch once compiled;
 *        JQ_JIT_SYNC      compile now (or wait for the worker) and load the
 *                         module; fails with the compiler log on error;
 *        JQ_JIT_QUERY     change nothing.
 *   *phase (optional) receives JQ_JIT_NONE / _COMPILING / _COMPILED /
 *   _LOADED / _FAILED.  A plan with a pass of more than 256 gates (a deep
 *   circuit on few wires) declines specialisation: _FAILED without an
 *   error, and the prebuilt kernels run it.
 * Without a call, plans on JQ_JIT_MIN_WIRES or more wires start in
 * JQ_JIT_ASYNC and smaller ones in JQ_JIT_OFF; the environment variable
 * JQ_JIT=0 / 1 overrides that default. */
#define JQ_JIT_OFF 0
#define JQ_JIT_ASYNC 1
#define JQ_JIT_SYNC 2
#define JQ_JIT_QUERY (-1)
#define JQ_JIT_NONE 0
#define JQ_JIT_COMPILING 1
#define JQ_JIT_COMPILED 2
#define JQ_JIT_LOADED 3
#define JQ_JIT_FAILED (-1)
#define JQ_JIT_MIN_WIRES 20
int jq_plan_jit(jq_plan* plan, int32_t mode, int32_t* phase);

/* Forward simulation of `batch` circuits sharing the plan's structure.
 *   angles   [batch][n_params] float64 (device)
 *   matrices [batch][n_matrix_elems] complex128 (device, may be NULL)
 *   psi0     initial vectors [batch or 1][2^n] in the plan dtype, or NULL for |0..0>
 *   state    output vectors [batch][2^n] (device)
 *   obs_*    observables (host arrays, n_obs entries: wire, d0, d1), kind JQ_OBS_*
 *   expvals  [batch][n_obs] float64 (device) or NULL
 * Replaces _run + measure (simulation.py:400-473, 735-857). */
int jq_sv_forward(jq_plan* plan, int32_t batch, const double* angles, const void* matrices,
                  const void* psi0, int32_t psi0_batched, void* state, int32_t obs_kind, int32_t n_obs,
                  const int32_t* obs_wires, const double* obs_d, double* expvals, void* stream);

/* Forward + adjoint-method gradient of  sum_o ct[b][o] <O_o>  per circuit.
 *   ct           [batch][n_obs] float64 (device)
 *   grad_angles  [batch][n_params] float64 (device, overwritten)
 *   grad_matrices[batch][n_matrix_elems] complex128 (device, JQ_MF_GRAD members; may be NULL)
 *   grad_psi0    [batch][2^n] plan dtype (device) or NULL: the initial-state cotangent
 *   state        [batch][2^n] workspace/output for the final state (device)
 * Replaces _adjoint_expval (simulation.py:546-655). */
int jq_sv_adjoint(jq_plan* plan, int32_t batch, const double* angles, const void* matrices,
                  const void* psi0, int32_t psi0_batched, void* state, int32_t obs_kind, int32_t n_obs,
                  const int32_t* obs_wires, const double* obs_d, const double* ct, double* expvals,
                  double* grad_angles, void* grad_matrices, void* grad_psi0, void* stream);

/* The adjoint in two calls, so that a caller that needs the forward's
 * outputs before it knows the cotangent (an autograd forward / backward
 * pair) runs the forward passes once:
 *   jq_sv_forward_ckpt  runs jq_sv_forward and keeps every pass output in the
 *                       plan; *token identifies that forward;
 *   jq_sv_backward      runs the adjoint sweep over those checkpoints.  The
 *                       caller passes the same batch, angles, matrices, psi0
 *                       and the untouched `state` of that forward.  If any
 *                       other call has used the plan since (the token no
 *                       longer matches), the forward is run again first, so
 *                       the result is always that of jq_sv_adjoint.
 * The cotangent is either ct (observables, as jq_sv_adjoint) or, when
 * lam_seed is not NULL, a caller-given cotangent of the final vector:
 * lam_seed [batch][2^n] in the plan dtype, JAX's convention (for a real loss
 * L, 2 dL/dpsi -- the conjugate of PyTorch's gradient).  lam_seed is used as
 * the running adjoint vector and overwritten.
 * Replaces the fwd/bwd pair of _adjoint_expval's custom_vjp
 * (simulation.py:586-654), and the reverse-mode autodiff the reference runs
 * through measure_state / measure_density for probs, state and density
 * outputs (simulation.py:658-732). */
int jq_sv_forward_ckpt(jq_plan* plan, int32_t batch, const double* angles, const void* matrices, const void* psi0,
                       int32_t psi0_batched, void* state, int32_t obs_kind, int32_t n_obs, const int32_t* obs_wires,
                       const double* obs_d, double* expvals, int64_t* token, void* stream);
int jq_sv_backward(jq_plan* plan, int64_t token, int32_t batch, const double* angles, const void* matrices,
                   const void* psi0, int32_t psi0_batched, void* state, int32_t obs_kind, int32_t n_obs,
                   const int32_t* obs_wires, const double* obs_d, const double* ct, void* lam_seed,
                   double* grad_angles, void* grad_matrices, void* grad_psi0, void* stream);

/* jq_sv_forward without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  The
 * plan's block matrices, expectation partials and observable table live in
 * a fixed per-device scratch array of JQ_NOALLOC_SCRATCH_BYTES; a call whose
 * workspace does not fit fails with JQ_ERR_ARG, as does a plan whose
 * program is not yet resident on the device (run any other entry point on
 * it once first).  Every call on a device shares that device's array:
 * calls that overlap on one GPU (other streams, other threads) race.  The
 * outputs equal jq_sv_forward's. */
#define JQ_NOALLOC_SCRATCH_BYTES (32u << 20)
int jq_sv_forward_noalloc(jq_plan* plan, int32_t batch, const double* angles, const void* matrices, const void* psi0,
                          int32_t psi0_batched, void* state, int32_t obs_kind, int32_t n_obs,
                          const int32_t* obs_wires, const double* obs_d, double* expvals, void* stream);

/* One group of a ragged batch: `batch` circuits sharing `plan`'s structure. */
typedef struct jq_group {
  jq_plan* plan;
  int32_t batch;
  int32_t psi0_batched;
  const double* angles;  /* [batch][n_params] float64 (device) */
  const void* matrices;  /* [batch][n_matrix_elems] complex128 (device) or NULL */
  const void* psi0;      /* [batch or 1][2^n] or NULL */
  const double* ct;      /* [batch][n_obs] float64 (device) */
  double* expvals;       /* [batch][n_obs] float64 (device) or NULL */
  double* grad_angles;   /* [batch][n_params] float64 (device) or NULL */
  void* state;           /* [batch][2^n] final states (device) */
} jq_group;

/* Forward + adjoint gradient of a ragged batch: several circuit structures
 * (e.g. uneven depths) of the same width and dtype, observables shared.
 * With the interpreter, every pass step is ONE launch over all circuits that
 * still have a pass at that step.  When every group's plan has its
 * plan-specialised kernels, each group runs its own adjoint on one of eight
 * side streams forked from `stream` and joined back to it with events (groups
 * sharing a plan share a side stream), so work enqueued on `stream` after
 * the call sees every output.  The reference runs one vmapped
 * program per structure group (script.py:406-434, the vmap in _vectorize
 * script.py:43-127); this call replaces that loop of execute() calls. */
int jq_sv_adjoint_groups(int32_t n_groups, const jq_group* groups, int32_t obs_kind, int32_t n_obs,
                         const int32_t* obs_wires, const double* obs_d, void* stream);

/* Reductions of a final state (measure_state, simulation.py:735-811). */
int jq_sv_probs(int32_t n_wires, int32_t batch, int32_t dtype, const void* state, void* probs, void* stream);
int jq_sv_outer(int32_t n_wires, int32_t batch, int32_t dtype, const void* state, void* rho, void* stream);
int jq_dm_diag(int32_t n_qubits, int32_t batch, int32_t dtype, const void* rho, void* probs, void* stream);

/* Expectation value of one k-qubit observable matrix (complex128 host array,
 * k <= 5) for each of `batch` states: Re <psi|O|psi>  (simulation.py:806-809),
 * or Re Tr(O rho) when density != 0 (simulation.py:850-852). */
int jq_expval_matrix(int32_t n_qubits, int32_t batch, int32_t dtype, int32_t density, const void* state,
                     int32_t k, const int32_t* wires, const double* matrix, double* out, void* stream);

/* Batched two-qubit pulse propagators in FP64 and the gradient of the
 * process infidelity.  Per parameter set the Hamiltonian is
 *   H(t) = w0/2 Z0 + w1/2 Z1 + J Z0 Z1
 *          + E(t) (cos phi X0 + sin phi Y0) + Q(t) (-sin phi X0 + cos phi Y0)
 * with E the lifted Gaussian of amplitude A and width sigma centred at T/2
 * and Q = -beta dE/dt (DRAG).  Magnus-2 midpoint steps, n_steps of them.
 *   params     [n_sets][6] float64 (device): A, sigma, beta, phi, w0, w1
 *   target     4x4 complex128 V, row-major, re/im interleaved (HOST memory)
 *   unitaries  [n_sets][4][4] complex128 (device) or NULL
 *   infidelity [n_sets] float64 (device): 1 - |Tr(V^dag U)|^2 / 16
 *   grad       [n_sets][4] float64 (device) or NULL: d infidelity / d(A, sigma, beta, phi)
 * Replaces the Magnus-2 solve of Evolution._evolve_parametrized
 * (evolution.py:194-263, 457-689) with the DRAG envelopes of
 * pulses.py:324-390, and unitary_cost_fn's process loss (qoc.py:261-343). */
#define JQ_ENV_DRAG_GAUSSIAN 1

typedef struct jq_pulse_desc {
  int32_t n_sets;
  int32_t n_steps;
  double t_total;    /* ns */
  double coupling_j; /* ZZ coupling (rad/ns) */
  int32_t envelope;  /* JQ_ENV_DRAG_GAUSSIAN */
  int32_t reserved;
} jq_pulse_desc;

int jq_pulse_propagate(const jq_pulse_desc* desc, const double* params, const double* target,
                       double* unitaries, double* infidelity, double* grad, void* stream);

/* Batched propagators of time-dependent Hamiltonians, in FP64:
 *   U_g = exp(-i A_{g,M-1}) ... exp(-i A_{g,1}) exp(-i A_{g,0}),
 *   A_{g,m} = sum_k coef[g][m][k] H_{g,k}   (H_k Hermitian, d = 2 or 4)
 * Every fixed-step scheme of the reference's Evolution is such a product
 * with host-side coefficients: Magnus-2 (one factor per step, coef = h f_k
 * at the midpoint), CFM4 (two factors per step, Gauss-point combinations)
 * and the single-term closed form exp(-i F H) (one factor, F the pulse
 * area).  Replaces Evolution._build_magnus_evolve_solver /
 * _build_closed_form_evolve_solver (evolution.py:194-263, 353-404) and the
 * batched resolve_pending solve (evolution.py:730-773).
 *   coef       [G][M][K] float64 (device)
 *   hmats      complex128, [G][K][d][d] if hmats_batched else [K][d][d] (device)
 *   unitaries  [G][d][d] complex128 (device, output)
 * jq_evolve_vjp: the gradient of a real loss L w.r.t. coef, given
 *   cot        [G][d][d] complex128: dL/dRe U + i dL/dIm U (PyTorch's convention)
 *   grad_coef  [G][M][K] float64 (output). */
typedef struct jq_evolve_desc {
  int32_t n_gates;
  int32_t dim;
  int32_t n_terms;
  int32_t n_factors;
  int32_t hmats_batched;
  int32_t reserved;
} jq_evolve_desc;

int jq_evolve(const jq_evolve_desc* desc, const double* coef, const double* hmats, double* unitaries, void* stream);
int jq_evolve_vjp(const jq_evolve_desc* desc, const double* coef, const double* hmats, const double* cot,
                  double* grad_coef, void* stream);

/* Kernel timing for measurement: when enabled, each run of consecutive
 * tile-pass launches of one kind (the forward passes of a call, its backward
 * passes) is bracketed by a pair of HIP events on its stream; jq_timing_read
 * returns the summed run durations and the number of launches they cover,
 * and clears them (which = 0 forward passes, 1 backward passes).  One event
 * pair per run, not per launch, keeps the measurement from stretching the
 * step.  For use from one host thread.  No reference counterpart. */
int jq_timing_enable(int32_t on);
int jq_timing_read(int32_t which, double* total_ms, int32_t* count);

/* Thread-local message of the last failing call. */
const char* jq_last_error(void);

/* Library version and the offload target it was built for. */
const char* jq_version(void);

#ifdef __cplusplus
}
#endif

#endif /* JAQSI_AMD_H */
