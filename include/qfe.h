/*
 * qfe.h -- C ABI of the MI355X state-vector engine behind qforte_amd.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes and a
 * thread-local message (qfe_last_error).  No C++ exception ever crosses this
 * boundary and no torch type appears in a signature.  Amplitudes are complex128
 * stored as interleaved (re, im) doubles, index bit q = qubit q, exactly the
 * layout of the reference Computer's std::vector<std::complex<double>>
 * (src/qforte/computer.h:176, qubit_basis.h:27-36), so upload/download are
 * plain copies.
 *
 * Each entry point names the reference interface it replaces.  Entry points
 * with no counterpart in the reference say so and name the callers they
 * serve.
 */
#ifndef QFE_H
#define QFE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define QFE_OK 0
#define QFE_ERR_ARG 1      /* invalid argument (bad handle, size mismatch, qubit out of range) */
#define QFE_ERR_DEVICE 2   /* HIP runtime error, or no GPU present                           */
#define QFE_ERR_NOMEM 3    /* device allocation failed                                        */
#define QFE_ERR_STATE 4    /* operation not valid in the current state                        */

typedef struct qfe_state_s* qfe_state;     /* one 2^n complex128 state resident in HBM      */
typedef struct qfe_program_s* qfe_program; /* compiled Pauli programs resident in HBM       */
typedef struct qfe_batch_s* qfe_batch;     /* serving engine (continuous batching)          */

typedef struct {
    double re, im;
} qfe_c64;

/* A gate in the reference's own form: target, control and the 4x4 matrix in
 * the basis 2*control_bit + target_bit, row-major (gate.h:52-154; a 1-qubit
 * gate has target == control and uses the top-left 2x2 block). */
typedef struct {
    int32_t target;
    int32_t control;
    qfe_c64 m[16];
} qfe_gate;

/* A Pauli operator in mask form: P|j> = phase * (-1)^popcount(j & z) |j ^ x>.
 * coeff multiplies P.  Compiled on the host from a canonical Pauli Circuit
 * (circuit.cc:107-169); phase is one of +1, -1, +i, -i. */
typedef struct {
    uint64_t x;
    uint64_t z;
    qfe_c64 phase;
    qfe_c64 coeff;
} qfe_pauli;

/* One element of an operation stream for qfe_apply_ops. */
#define QFE_OP_GATE 0 /* a gate (uses .gate)                                                  */
#define QFE_OP_PEXP 1 /* exp(i * angle * P) for a Hermitian P (uses .pauli, .angle)            */
typedef struct {
    int32_t kind;
    int32_t pad;
    double angle;
    qfe_gate gate;
    qfe_pauli pauli;
} qfe_op;

/* ---- library ------------------------------------------------------------ */
/* No reference counterpart (the reference has no device): error reporting. */
const char* qfe_last_error(void);
/* Select the GPU and create the library stream.  Called implicitly by the
 * first state creation; no reference counterpart. */
int qfe_init(int device);
/* Number of visible GPUs (0 on a host without one); no reference counterpart. */
int qfe_device_count(int* count);
/* The hipStream_t every kernel of the library runs on (for event timing by
 * callers); no reference counterpart. */
int qfe_get_stream(void** stream);
/* Block until all queued work finished. */
int qfe_synchronize(void);
/* Fusion / launch tuning: max tile qubits (k), low contiguous qubits (L) and
 * max ops per fused pass.  Zero leaves a value unchanged; -1 (any argument)
 * restores the default, a shape chosen by state size ((8, 4, 64) up to 14
 * qubits, (10, 4, 32) at 15-20, (11, 4, 24) above; register-orbit passes
 * take at most 24 ops), for which qfe_get_tuning reports -1 -1 -1.  No
 * reference counterpart (autotuner hook). */
int qfe_set_tuning(int tile_qubits, int low_qubits, int max_ops_per_pass);
int qfe_get_tuning(int* tile_qubits, int* low_qubits, int* max_ops_per_pass);
/* Largest register orbit of the planner, 1..7 dimensions (5, 6 and 7 spread one
 * orbit over 2, 4 or 8 lanes of a wave); no reference counterpart (autotuner hook). */
int qfe_set_pass_tuning(int orbit_max_dims);
int qfe_get_pass_tuning(int* orbit_max_dims);
/* Largest number of determinant pairs for which a sector template stores one
 * 8-byte pair word per pair (templates above it rotate from the compact
 * string lists); applies to templates created afterwards.  Default 2^25.  No
 * reference counterpart (memory / speed trade-off of the sector engine). */
int qfe_set_sector_word_limit(uint64_t max_pair_words);
/* HIP events on the library stream, for timing kernels from the host
 * (no reference counterpart). */
int qfe_event_create(void** event);
int qfe_event_record(void* event);
int qfe_event_elapsed_ms(void* start, void* stop, float* ms);
int qfe_event_destroy(void* event);
/* Counters of launched passes since the last reset (fused passes, single-op
 * passes, expectation passes) and the algorithmic bytes they moved. */
int qfe_get_counters(uint64_t* fused_passes, uint64_t* single_passes, uint64_t* reduce_passes, double* bytes);
/* The same counters split by kind: out[0..count) of {fused passes, single-op
 * passes, expectation passes, LDS-tile passes, register-orbit passes,
 * lane-orbit passes, passes with non-temporal loads and stores}; no
 * reference counterpart (tests and benchmarks read which kernels ran). */
int qfe_get_pass_counts(uint64_t* out, uint64_t count);
int qfe_reset_counters(void);

/* ---- state (Computer storage, computer.cc:28-54, computer.h:28,112,132) -- */
/* Computer::Computer(nqubit): allocate and set |0...0> (computer.cc:28-38). */
int qfe_state_create(int nqubit, qfe_state* out);
int qfe_state_destroy(qfe_state s);
int qfe_state_nqubit(qfe_state s, int* nqubit);
/* Computer(const Computer&) deep copy (computer.h:28): device-to-device. */
int qfe_state_copy(qfe_state dst, qfe_state src);
/* set_coeff_vec / set_coeff_vec_from_numpy (computer.h:132, bindings.cc:176-182). */
int qfe_state_upload(qfe_state s, const qfe_c64* host, uint64_t namps);
/* get_coeff_vec (computer.h:112). */
int qfe_state_download(qfe_state s, qfe_c64* host, uint64_t namps);
/* reset (computer.cc:51-54). */
int qfe_state_reset(qfe_state s);
/* null_state (computer.cc:49). */
int qfe_state_zero(qfe_state s);
/* set_state: zero, then scatter (index, value) pairs (computer.cc:42-47). */
int qfe_state_set_amps(qfe_state s, const uint64_t* idx, const qfe_c64* val, uint64_t count);
/* coeff(QubitBasis) (computer.cc:40). */
int qfe_state_get_amp(qfe_state s, uint64_t idx, qfe_c64* out);
/* apply_constant (computer.cc:130-133). */
int qfe_state_scale(qfe_state s, qfe_c64 a);
/* y += a * x; the accumulate step of apply_operator (computer.cc:84-86). */
int qfe_state_axpy(qfe_state y, qfe_c64 a, qfe_state x);
/* dot(qc1, qc2) = sum conj(a_i) b_i (computer.cc:975-980), deterministic order. */
int qfe_state_dot(qfe_state a, qfe_state b, qfe_c64* out);
/* operator== (computer.cc:968-973): exact equality of every amplitude. */
int qfe_state_equal(qfe_state a, qfe_state b, int* equal);
/* Probability scan sum |c_i|^2 (norm squared); no direct reference counterpart,
 * used by norm checks. */
int qfe_state_norm2(qfe_state s, double* out);
/* Draw nsamples computational-basis indices from |psi|^2: the probability
 * scan and std::discrete_distribution draw of Computer::measure_circuit,
 * measure_z_readouts_fast and measure_readouts (computer.cc:166-188,
 * 195-210, 247-262; the basis change before it is an ordinary apply on a
 * copy).  A counter-based generator seeded by `seed` draws the uniforms, so a
 * seed repeats its indices (the reference seeds from std::random_device).
 * out: nsamples indices in host memory; at most 2^32 per call. */
int qfe_sample(qfe_state s, uint64_t nsamples, uint64_t seed, uint64_t* out);

/* ---- gates and Pauli exponentials --------------------------------------- */
/* apply_gate / apply_circuit (computer.cc:93-114, 340-751).  Gates are fused
 * into multi-gate passes where that is exact. */
int qfe_apply_gates(qfe_state s, const qfe_gate* gates, uint64_t count);
/* A mixed stream of gates and Pauli exponentials, applied in order.  A
 * QFE_OP_PEXP replaces the gate sequence exponentiate_pauli_string builds
 * for exp(i*angle*P) (exponentiate.py:9-87) by one pass. */
int qfe_apply_ops(qfe_state s, const qfe_op* ops, uint64_t count);

/* ---- Pauli sums (computer.cc:80-91, 753-806) ----------------------------- */
/* direct_op_exp_val for an operator of Pauli terms: <psi| sum_k coeff_k P_k |psi>. */
int qfe_expval(qfe_state s, const qfe_pauli* terms, uint64_t count, qfe_c64* out);
/* qfe_expval with no allocation: no reference counterpart; for callers that
 * cannot allocate device memory (code under hipGraph capture, tight serving
 * loops, hosts that forbid hipMalloc).  The group and term tables, partial
 * sums and per-group results live in one fixed 6 MiB device array per GPU
 * (resolved once by qfe_init), so at most 65536 terms per call (QFE_ERR_ARG
 * beyond).  Result equal to qfe_expval's.  Calls that overlap on one GPU
 * race on that array; every call synchronizes the library stream before it
 * returns. */
int qfe_expval_noalloc(qfe_state s, const qfe_pauli* terms, uint64_t count, qfe_c64* out);
/* Several operators at once (direct_oppl_exp_val, computer.cc:773-806):
 * op_offsets[k]..op_offsets[k+1] delimit operator k; out[k] = <psi|O_k|psi>. */
int qfe_expval_many(qfe_state s, const qfe_pauli* terms, const uint64_t* op_offsets, uint64_t nops,
                    qfe_c64* out);
/* apply_operator: out = sum_k coeff_k P_k in (computer.cc:80-91).  out != in. */
int qfe_apply_pauli_sum(qfe_state in, const qfe_pauli* terms, uint64_t count, qfe_state out);
/* <bra| sum_k coeff_k P_k |ket>: the np.vdot(sigma, K psi) of
 * measure_gradient / measure_gradient3 (uccvqeabc.py:159-162, 242-244,
 * 274-278) without forming K psi. */
int qfe_cross_expval(qfe_state bra, qfe_state ket, const qfe_pauli* terms, uint64_t count, qfe_c64* out);
/* Several operators in one launch: out[k] = <bra|O_k|ket>, operators
 * delimited by op_offsets (the per-operator vdot loops of QITE's build_b and
 * build_S, qite.py:256-326, and of measure_gradient3). */
int qfe_cross_expval_many(qfe_state bra, qfe_state ket, const qfe_pauli* terms, const uint64_t* op_offsets,
                          uint64_t nops, qfe_c64* out);

/* ---- compiled UCC programs (no single reference counterpart) ------------
 * A program holds, resident in HBM, the exponential sequence of a
 * disentangled UCC ansatz (ansatz.py:25-82 + trotterization.py:10-57) and a
 * Hamiltonian (Molecule.hamiltonian).  Amplitudes are runtime inputs, so one
 * compile serves every energy and gradient call of an optimisation
 * (energy_feval, algorithm.py:458-474; measure_gradient, uccvqeabc.py:122-251).
 *
 * Generators: generator k has terms gen_offsets[k]..gen_offsets[k+1] (each
 * term's coeff is the purely imaginary JW coefficient); the ansatz applies
 * exp(t[param[k]] * scale[k] * sum_j coeff_j P_j) for k = 0, 1, ... with the
 * terms in order, as trotterize does. */
int qfe_program_create(int nqubit, uint64_t ref_index, const qfe_pauli* gen_terms, const uint64_t* gen_offsets,
                       uint64_t ngen, const int32_t* gen_param, const double* gen_scale, uint64_t nparam,
                       const qfe_pauli* ham_terms, uint64_t nham, qfe_program* out);
int qfe_program_destroy(qfe_program p);
/* E(t) = <ref|U(t)^dag H U(t)|ref> (energy_feval, algorithm.py:458-474). */
int qfe_program_energy(qfe_program p, const double* params, double* energy);
/* Prepare U(t)|ref> into a state (build_Uvqc + apply_circuit). */
int qfe_program_prepare(qfe_program p, const double* params, qfe_state out);
/* qfe_program_energy through a captured hipGraph (no reference counterpart;
 * for optimizer loops that evaluate one ansatz many times).  The first call
 * plans the ansatz once with device-side angles and captures parameter upload,
 * angle update, reference state, exponential passes, expectation value and
 * result download on the library stream; every call writes the amplitudes and
 * replays the graph.  The plan is the one the tuning in force at that first
 * call gives.  Equals qfe_program_energy to rounding (1e-12). */
int qfe_program_energy_graph(qfe_program p, const double* params, double* energy);
/* Adjoint gradient dE/dt_k (measure_gradient, uccvqeabc.py:122-251): one
 * forward sweep, sigma = H psi, then a reverse sweep un-applying each
 * exponential on psi and sigma.  grad has nparam entries; energy may be NULL. */
int qfe_program_gradient(qfe_program p, const double* params, double* grad, double* energy);
/* PQE residual amplitudes (UCCNPQE.get_residual_vector, uccnpqe.py:249-292:
 * apply_circuit(U), apply_operator(H), apply_circuit(U.adjoint()), then
 * coeff at each excited determinant): out[i] = <idx[i]| U(t)^dag H U(t) |ref>,
 * idx[i] a basis index of the register.  The caller multiplies by the
 * determinant's phase. */
int qfe_program_residuals(qfe_program p, const double* params, const uint64_t* idx, uint64_t count, qfe_c64* out);
/* ADAPT pool gradients 2 Re <H psi | A_k psi> for many operators A_k at the
 * state U(t)|ref> (measure_gradient3, uccvqeabc.py:253-283).  Operators are
 * delimited by op_offsets; processed in batches of up to `batch` operators
 * per launch. */
int qfe_program_pool_gradients(qfe_program p, const double* params, const qfe_pauli* pool_terms,
                               const uint64_t* op_offsets, uint64_t nops, uint64_t batch, double* grads);

/* ---- determinant-sector templates --------------------------------------
 * A template compiles one size class of UCC energy evaluations: a
 * disentangled ansatz on the determinant ref_index, and a Hamiltonian whose
 * coefficients are linear in per-request inputs.  Replaces, per request,
 * UCC.ansatz_circuit + trotterize + measure_energy (ansatz.py:25-82,
 * algorithm.py:271-334, 458-480) with the same mathematics.
 *
 * The generators (gen_terms split by gen_offsets, applied in order, each
 * as exp(scale * params[param] * sum_j coeff_j * phase_j * P_j), term by
 * term) must conserve the particle numbers on even bits (alpha) and odd
 * bits (beta): every run of consecutive terms with one X mask and mutually
 * commuting strings must map sector determinants to sector determinants.
 * The state is then kept on the determinants of the reference's
 * (N_alpha, N_beta) sector only, and each run is applied exactly as the
 * product of its commuting exponentials: a rotation of determinant pairs.
 * QFE_ERR_ARG if a run does not conserve the sector.
 *
 * Hamiltonian term k has masks (x, z) and phase from ham_terms[k]; its
 * coefficient is phase_k * sum_{e = map_offsets[k]}^{map_offsets[k+1]-1}
 * map_value[e] * inputs[map_index[e]] with per-request real inputs, or
 * phase_k * ham_terms[k].coeff when map_offsets is NULL (ninputs = 0). */
typedef struct qfe_template_s* qfe_template;
int qfe_template_create(int nqubit, uint64_t ref_index, const qfe_pauli* gen_terms, const uint64_t* gen_offsets,
                        uint64_t ngen, const int32_t* gen_param, const double* gen_scale, uint64_t nparam,
                        const qfe_pauli* ham_terms, uint64_t nham, const uint64_t* map_offsets,
                        const uint32_t* map_index, const qfe_c64* map_value, uint64_t ninputs, qfe_template* out);
int qfe_template_destroy(qfe_template t);
/* Sector dimension and device bytes one request of this template occupies. */
int qfe_template_info(qfe_template t, uint64_t* sector_dim, uint64_t* request_bytes);
/* Host-side audit of a captured single-request sequence (no reference
 * counterpart; a diagnostic for callers and tests).  Templates whose state
 * does not fit one workgroup's LDS replay their energy (gradient = 0) and
 * gradient (gradient = 1) sequences as hipGraphs; every capture is checked
 * before it is kept: one chain of nodes, each node the expected memcpy or
 * kernel, grids non-zero, blocks of 1..1024 threads, arguments and copy
 * extents equal to the addresses the template owns.  Builds the graph if
 * needed and writes out4 = {nodes, kernel nodes, kernel nodes whose arguments
 * were compared, memcpy nodes whose parameters were compared}; all zero for
 * templates that launch without a graph. */
int qfe_template_graph_audit(qfe_template t, int gradient, uint64_t* out4);
/* One request in one launch (no batching): E = <ref|U^dag H U|ref>. */
int qfe_template_energy(qfe_template t, const double* params, const double* inputs, double* energy);
/* U|ref> scattered into a full 2^n state (for checks against the full-space path). */
int qfe_template_prepare(qfe_template t, const double* params, qfe_state out);
/* Adjoint gradient dE/dt on the sector (uccvqeabc.py:122-251 recursion):
 * sigma = H psi, then a reverse sweep un-applying each generator on psi and
 * sigma; grad[p] = sum over generators k with param p of 2 scale_k
 * Re<sigma_k|K_k|psi_k>.  energy (may be NULL) receives <psi|H|psi>. */
int qfe_template_gradient(qfe_template t, const double* params, const double* inputs, double* grad, double* energy);
/* PQE residual amplitudes on the sector (as qfe_program_residuals:
 * uccnpqe.py:249-292); every idx[i] must lie in the template's sector. */
int qfe_template_residuals(qfe_template t, const double* params, const double* inputs, const uint64_t* idx,
                           uint64_t count, qfe_c64* out);
/* ADAPT pool gradients 2 Re<H psi|A_k psi> (uccvqeabc.py:253-283) for nops
 * anti-Hermitian operators given as Pauli sums (pool_terms split by
 * op_offsets); `batch` operators per launch (0: 64). */
int qfe_template_pool_gradients(qfe_template t, const double* params, const double* inputs, const qfe_pauli* pool_terms,
                                const uint64_t* op_offsets, uint64_t nops, uint64_t batch, double* grads);

/* ---- serving: continuous batching of UCC energy and gradient requests ----
 * No reference counterpart: the reference evaluates one energy per Python
 * call.  Requests (a template, amplitudes and Hamiltonian inputs) queue on
 * the host; each engine step admits queued requests into a launch while the
 * launch's device footprint stays within its share of hbm_budget_bytes
 * (at most 256 MB per launch, the size of the last-level cache) and max_active, runs every admitted request (of any template and size) in one
 * kernel, one workgroup per request, on one of a few streams, so launches
 * overlap on the device while later arrivals queue for the next launch. */
int qfe_batch_create(uint64_t hbm_budget_bytes, int max_active, qfe_batch* out);
int qfe_batch_destroy(qfe_batch b);
/* Queue a request; params and inputs are copied.  The template must outlive it. */
int qfe_batch_submit(qfe_batch b, qfe_template t, const double* params, const double* inputs, uint64_t* request_id);
/* Queue an energy + adjoint gradient request (the qfe_template_gradient
 * computation; uccvqeabc.py:122-251 for the reference's per-call form).  Its
 * energy is returned by step/poll like any other; once it has been returned
 * the gradient (nparam values) is fetched once with qfe_batch_take_gradient.
 * A launch holds requests of one kind, taken from the kind whose oldest
 * queued request arrived first. */
int qfe_batch_submit_gradient(qfe_batch b, qfe_template t, const double* params, const double* inputs,
                              uint64_t* request_id);
int qfe_batch_take_gradient(qfe_batch b, uint64_t request_id, double* grad, uint64_t nparam);
/* Launch what can be admitted; return completed results (ids, energies, up to
 * cap).  Blocks for the oldest launch only when nothing else can progress. */
int qfe_batch_step(qfe_batch b, uint64_t* done_ids, double* done_energies, uint64_t cap, uint64_t* ndone);
/* Non-blocking form: launch what can be admitted and return what is done. */
int qfe_batch_poll(qfe_batch b, uint64_t* done_ids, double* done_energies, uint64_t cap, uint64_t* ndone);
int qfe_batch_pending(qfe_batch b, uint64_t* queued, uint64_t* active);

#ifdef __cplusplus
}
#endif

#endif /* QFE_H */
